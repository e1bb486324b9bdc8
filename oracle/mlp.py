"""CPU restatement of the learned proposal's network and features (numpy, float64).

TEST INFRASTRUCTURE ONLY: the checker of the library's FP32 training kernels and of its
device-side features (sunhubbardmy_amd/csrc/learn_kernels.hip).  The reference has no learned
proposal, so nothing here restates reference code; each function states the definition it
computes, and tests/test_oracle_learn.py pins it against torch autograd, finite differences
and brute-force loops.

Weights layout (float32 vector): W1[128][4], b1[128], W2[128][128], b2[128], w3[128], b3.
Network: g(x) = b3 + w3 . tanh(b2 + W2 tanh(b1 + W1 x)); x = (f - xmu) / xsd.
Loss: mean over samples of (g(x) - (E/Ns - ymu) / ysd)^2.
"""
import numpy as np

H = 128
NW = H * H + 7 * H + 1


def unpack(w):
    w = np.asarray(w, np.float64)
    o = 0
    out = []
    for shape in ((H, 4), (H,), (H, H), (H,), (H,), (1,)):
        size = int(np.prod(shape))
        out.append(w[o:o + size].reshape(shape))
        o += size
    return out


def features(conf, nbr):
    """f0..f3 of an HS field conf[ltrot, N] (values 1, 2) with neighbour lists nbr[N, z]."""
    s = 2 * np.asarray(conf, np.int64) - 3
    lt, n = s.shape
    nb2 = nbr.shape[0] * nbr.shape[1]
    s1 = np.roll(s, -1, axis=0)                      # s(t+1), periodic
    nn = s[:, nbr].sum(axis=2)                       # sum over neighbours, same slice
    nn1 = s1[:, nbr].sum(axis=2)
    ns = float(n * lt)
    return np.array([s.sum() / ns, (s * nn).sum() / (nb2 * lt), (s * s1).sum() / ns, (s * nn1).sum() / (nb2 * lt)])


def forward(w, norm, f):
    """ymu + ysd g(x) for features f[n, 4]: E_eff / Ns."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    norm = np.asarray(norm, np.float64)
    x = (np.asarray(f, np.float64).reshape(-1, 4) - norm[:4]) / norm[4:8]
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    g = h2 @ w3 + b3[0]
    return norm[8] + norm[9] * g


def loss_grad(w, norm, samples):
    """Mean squared error of g against the normalised energies, and its gradient (weights layout)."""
    W1, b1, W2, b2, w3, b3 = unpack(w)
    norm = np.asarray(norm, np.float64)
    samples = np.asarray(samples, np.float64).reshape(-1, 5)
    n = samples.shape[0]
    x = (samples[:, :4] - norm[:4]) / norm[4:8]
    y = (samples[:, 4] - norm[8]) / norm[9]
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    g = h2 @ w3 + b3[0]
    e = g - y
    loss = float(np.mean(e * e))
    dg = 2.0 * e / n
    gw3 = h2.T @ dg
    gb3 = np.array([dg.sum()])
    da2 = np.outer(dg, w3) * (1.0 - h2 * h2)
    gW2 = da2.T @ h1
    gb2 = da2.sum(axis=0)
    da1 = (da2 @ W2) * (1.0 - h1 * h1)
    gW1 = da1.T @ x
    gb1 = da1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, gb3])
    return loss, grad
