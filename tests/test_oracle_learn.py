"""Pins of the learned proposal's oracle (oracle/mlp.py and the block move of oracle/dqmc_oracle.c).

The reference has no learned proposal, so these pins come from outside the oracle:

  features            brute-force loops over sites, slices and neighbours
  loss and gradient   torch autograd in float64; central finite differences
  block move          exact enumeration of every HS field of a 4-site, 2-slice system: with
                      block moves only (no local update) and a deliberately strong random
                      effective action, the chain must still sample |W(C)| exactly, which fails
                      if the determinant ratio, the flipped set, the incremental feature sums
                      or the effective-action correction is wrong; the G update of an accepted
                      block is pinned by the wrap error against the freshly stabilised G
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import mlp
from sunhubbardmy_amd.model import Model, lattice_tables, make_tables


def _rand_weights(seed, scale=1.0):
    r = np.random.default_rng(seed)
    return (r.standard_normal(mlp.NW) * scale * np.sqrt(2.0 / 132)).astype(np.float32)


def test_features_brute_force():
    r = np.random.default_rng(3)
    for lat, la, lb in (("square", 4, 4), ("honeycomb", 3, 2), ("square", 2, 2)):
        nbr = lattice_tables(Model(lattice=lat, la=la, lb=lb))["nnlist"]
        n, z = nbr.shape
        lt = 5
        conf = r.integers(1, 3, size=(lt, n))
        s = 2 * conf - 3
        t = [0.0, 0.0, 0.0, 0.0]
        for tau in range(lt):
            for i in range(n):
                t[0] += s[tau, i]
                t[2] += s[tau, i] * s[(tau + 1) % lt, i]
                for k in range(z):
                    j = nbr[i, k]
                    t[1] += s[tau, i] * s[tau, j]
                    t[3] += s[tau, i] * s[(tau + 1) % lt, j]
        want = [t[0] / (n * lt), t[1] / (n * z * lt), t[2] / (n * lt), t[3] / (n * z * lt)]
        assert np.allclose(mlp.features(conf, nbr), want, rtol=0, atol=1e-15)


def _torch_loss(w, norm, x):
    W1, b1, W2, b2, w3, b3 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in mlp.unpack(w)]
    norm = torch.tensor(norm, dtype=torch.float64)
    xs = torch.tensor(x, dtype=torch.float64)
    xi = (xs[:, :4] - norm[:4]) / norm[4:8]
    y = (xs[:, 4] - norm[8]) / norm[9]
    g = torch.tanh(torch.tanh(xi @ W1.T + b1) @ W2.T + b2) @ w3 + b3[0]
    loss = ((g - y) ** 2).mean()
    loss.backward()
    loss = loss.detach()
    grad = torch.cat([W1.grad.reshape(-1), b1.grad, W2.grad.reshape(-1), b2.grad, w3.grad, b3.grad]).numpy()
    return float(loss), grad


def _samples(seed, n):
    r = np.random.default_rng(seed)
    x = np.zeros((n, 5))
    x[:, :4] = r.normal([0.01, 0.2, 0.6, 0.05], [0.01, 0.03, 0.05, 0.02], size=(n, 4))
    x[:, 4] = -0.3 + 0.02 * r.standard_normal(n)
    norm = np.array([0.01, 0.2, 0.6, 0.05, 0.01, 0.03, 0.05, 0.02, -0.3, 0.02])
    return x, norm


def test_loss_grad_matches_autograd():
    w = _rand_weights(1)
    x, norm = _samples(2, 37)
    l0, g0 = mlp.loss_grad(w, norm, x)
    l1, g1 = _torch_loss(w, norm, x)
    assert abs(l0 - l1) < 1e-12 * max(1.0, l1)
    assert np.abs(g0 - g1).max() < 1e-12 * max(1.0, np.abs(g1).max())


def test_loss_grad_finite_differences():
    w = _rand_weights(4).astype(np.float64)
    x, norm = _samples(5, 11)
    _, g = mlp.loss_grad(w, norm, x)
    r = np.random.default_rng(6)
    for k in list(r.integers(0, mlp.NW, 12)) + [mlp.NW - 1, 0, 4 * 128, 5 * 128]:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mlp.loss_grad(wp, norm, x)[0] - mlp.loss_grad(wm, norm, x)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-7 * max(1.0, abs(g[k])), k


def test_forward_is_loss_consistent():
    w = _rand_weights(7)
    x, norm = _samples(8, 5)
    e = mlp.forward(w, norm, x[:, :4])
    # loss of samples whose target is the prediction itself is zero
    y = x.copy()
    y[:, 4] = e
    assert mlp.loss_grad(w, norm, y)[0] < 1e-26


KW2 = dict(lattice="square", la=2, lb=2, beta=1.0, dtau=0.5, rhub=6.0, kmode="checkerboard", nwrap=1,
           update="delay", nublock=2)


def _enumerate(kw):
    """|W(C)| of every field of the 2x2, 2-slice system, and each field's features."""
    from test_oracle import _b_product_checkerboard
    t = make_tables(Model(**kw))
    n, lt = t["ndim"], t["ltrot"]
    hs = t["hs"]
    nbr = t["lat"]["nnlist"]
    ws, fs = [], []
    for code in range(2 ** (n * lt)):
        conf = np.array([(code >> k) & 1 for k in range(n * lt)], np.int32).reshape(lt, n) + 1
        B = _b_product_checkerboard(t, conf, hs["bmat"])
        w = np.linalg.det(np.eye(n) + B) ** t["model"].nflr * np.prod(hs["phase"][conf - 1])
        ws.append(abs(w.real))
        fs.append(mlp.features(conf, nbr))
    ws = np.array(ws)
    return ws / ws.sum(), np.array(fs), nbr


def _code(conf):
    bits = (np.asarray(conf).reshape(-1) - 1).astype(np.int64)
    return int((bits << np.arange(bits.size)).sum())


@pytest.fixture(scope="module")
def enum2():
    return _enumerate(KW2)


def _strong_network():
    # a random network whose effective action changes by about one unit per spin flip of the
    # 8-spin field, so that the inner chain alone would sample a clearly different distribution
    w = _rand_weights(11, scale=1.0)
    norm = np.array([0.0, 0.0, 0.0, 0.0, 0.3, 0.3, 0.3, 0.3, 0.0, 0.3])
    return w, norm


def test_block_moves_sample_the_exact_distribution(enum2):
    p, fs, nbr = enum2
    w, norm = _strong_network()
    # what the inner chain alone would sample: exp(-E_eff) reweighting of p
    e = mlp.forward(w, norm, fs) * 8
    q = p * np.exp(-(e - e.min()))
    q /= q.sum()
    f_exact = p @ fs
    f_biased = q @ fs
    assert np.abs(f_exact - f_biased).max() > 0.05      # the test can tell the two apart
    hist = np.zeros(p.size)
    means = []
    for c in range(6):
        o = oracle.OracleChain(seed=oracle.chain_seed(c), **KW2)
        o.init_conf()
        o.sweep_start()
        o.set_proposal(nbr, 6, w, norm, local_updates=False)
        for _ in range(100):
            o.sweep(lupdate=True, lmeasure=False)
        fsum = np.zeros(4)
        cnt = 0
        for _ in range(3000):
            for part in ("b0", "0b"):
                getattr(o, "sweep_" + part)(True, False)
                k = _code(o.conf)
                hist[k] += 1
                fsum += fs[k]
                cnt += 1
        means.append(fsum / cnt)
        d = o.diag()
        assert d["max_wrap_error"] < 1e-10
        st = o.proposal_stats()
        assert 0 < st["accepted"] < st["moves"]
    means = np.array(means)
    mean = means.mean(0)
    err = means.std(0, ddof=1) / np.sqrt(len(means))
    assert np.all(np.abs(mean - f_exact) < 5 * err + 2e-3), (mean, f_exact, err)
    hist /= hist.sum()
    tv = 0.5 * np.abs(hist - p).sum()
    tv_biased = 0.5 * np.abs(hist - q).sum()
    assert tv < 0.07 and tv_biased > 2 * tv, (tv, tv_biased)


def test_block_moves_with_local_updates_keep_g_exact():
    kw = dict(lattice="square", la=4, lb=4, beta=2.0, dtau=0.1, rhub=4.0, kmode="checkerboard", nwrap=5,
              update="delay", nublock=4)
    nbr = lattice_tables(Model(**kw))["nnlist"]
    w, norm = _strong_network()
    norm = norm.copy()
    norm[9] = 0.01
    o = oracle.OracleChain(seed=oracle.chain_seed(2), **kw)
    o.init_conf()
    o.sweep_start()
    o.set_proposal(nbr, 16, w, norm, local_updates=True)
    for _ in range(2):
        o.sweep()
    st = o.proposal_stats()
    assert st["moves"] == 2 * 2 * o.ltrot and st["accepted"] > 0
    d = o.diag()
    assert d["max_wrap_error"] < 1e-10
