"""GPU tests of the learned block proposals (include/sunhip_learn.h) against the CPU oracle.

  features            device features of the HS field == oracle/mlp.features (exact integers)
  network             FP32 loss, gradient and prediction of the library vs the float64 oracle
                      (relative 2e-4: FP32 accumulation over 128-wide layers)
  training            Adam on samples of a known smooth function drives the loss down
  block move parity   with identical seeds, weights and normalisation the HIP sweep with block
                      proposals and the oracle's restatement produce bit-identical HS fields
                      after every pass, G within 1e-10 and identical proposal statistics
  detailed balance    block moves alone sample |W(C)| of a 4-site, 2-slice system exactly
  what training gains trained effective action vs an untrained one on the same chains
"""
import sys

import numpy as np
import pytest

import oracle
from oracle import mlp
from sunhubbardmy_amd import FTWalkers, Model, chain_seed
from sunhubbardmy_amd.learned import EffectiveAction
from sunhubbardmy_amd.model import lattice_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _rand_weights(seed, scale=1.0):
    r = np.random.default_rng(seed)
    return (r.standard_normal(mlp.NW) * scale * np.sqrt(2.0 / 132)).astype(np.float32)


def _samples(seed, n):
    r = np.random.default_rng(seed)
    x = np.zeros((n, 5))
    x[:, :4] = r.normal([0.01, 0.2, 0.6, 0.05], [0.01, 0.03, 0.05, 0.02], size=(n, 4))
    x[:, 4] = -0.3 + 0.02 * r.standard_normal(n)
    norm = np.array([0.01, 0.2, 0.6, 0.05, 0.01, 0.03, 0.05, 0.02, -0.3, 0.02])
    return x, norm


def test_features_match_oracle():
    kw = dict(lattice="square", la=4, lb=4, beta=1.0, dtau=0.1, rhub=4.0, nwrap=5, nublock=4)
    w = FTWalkers([Model(seed=chain_seed(c), **kw) for c in range(3)] +
                  [Model(seed=chain_seed(9), **dict(kw, la=2, lb=2))])
    say("features: 4 chains")
    w.set_neighbours()
    w.init_conf()
    w.sweep_start_0b()
    f = w.features()
    for c in range(w.nchains):
        nbr = w.tables[c]["lat"]["nnlist"]
        assert np.array_equal(f[c, :4], mlp.features(w.get_conf(c), nbr)), c
        d = w.diag(c)
        n, lt = w.tables[c]["ndim"], w.tables[c]["ltrot"]
        assert abs(f[c, 4] - (-2.0 * d["logweightf"].real / (n * lt))) < 1e-12
    w.close()


def test_network_matches_oracle():
    net = EffectiveAction(capacity=4096, seed=3)
    wts = _rand_weights(1)
    net.set_weights(wts)
    x, norm = _samples(2, 301)
    net.set_norm(norm)
    say("network: loss/grad over", len(x))
    loss, g = net.loss_grad(x)
    l0, g0 = mlp.loss_grad(wts, norm, x)
    assert abs(loss - l0) < 2e-4 * l0
    assert relmax(g.astype(np.float64), g0) < 2e-4
    e = net.predict(x[:, :4])
    e0 = mlp.forward(wts, norm, x[:, :4])
    assert relmax(e - norm[8], e0 - norm[8]) < 2e-5
    # Glorot initialisation of a fresh network: zero biases, bounded weights
    w0 = EffectiveAction(capacity=16, seed=5).weights()
    v = dict(zip(("W1", "b1", "W2", "b2", "w3", "b3"), mlp.unpack(w0)))
    assert np.all(v["b1"] == 0) and np.all(v["b2"] == 0) and v["b3"][0] == 0
    assert np.abs(v["W2"]).max() <= np.sqrt(6.0 / 256) and v["W2"].std() > 0.05
    net.close()


def test_training_fits_a_smooth_function():
    r = np.random.default_rng(7)
    n = 4000
    x = np.zeros((n, 5))
    x[:, :4] = r.normal(0.0, 1.0, size=(n, 4))
    x[:, 4] = np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2
    net = EffectiveAction(capacity=n, seed=11)
    net.push_samples(x)
    assert net.num_samples() == n
    assert np.array_equal(net.samples(), x)
    say("training: 40 epochs over", n)
    hist = net.train(epochs=40, batch=128, lr=2e-3)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.05 * hist[0] and hist[-1] < 0.03, hist
    # the norm was taken from the samples
    nm = net.norm()
    assert abs(nm[8] - x[:, 4].mean()) < 1e-12 and abs(nm[9] - x[:, 4].std()) < 1e-12
    # prediction agrees with the oracle forward of the trained weights
    e = net.predict(x[:50, :4])
    e0 = mlp.forward(net.weights(), nm, x[:50, :4])
    assert np.abs(e - e0).max() < 1e-4
    net.close()


PAR = {
    "square4_local": (dict(lattice="square", la=4, lb=4, beta=1.0, dtau=0.1, rhub=4.0, kmode="checkerboard",
                           nwrap=5, update="delay", nublock=4), 16, True),
    "square4_blocks_only": (dict(lattice="square", la=4, lb=4, beta=1.0, dtau=0.1, rhub=4.0, kmode="checkerboard",
                                 nwrap=5, update="delay", nublock=4), 24, False),
    "honeycomb3_dense": (dict(lattice="honeycomb", la=3, lb=3, beta=1.2, dtau=0.1, rhub=3.0, kmode="dense",
                              nwrap=4, update="delay", nublock=8), 64, True),
}


@pytest.mark.parametrize("name", list(PAR))
def test_block_move_parity(name):
    kw, nprop, local = PAR[name]
    seed = chain_seed(2)
    wts = _rand_weights(13, scale=1.0)
    norm = np.array([0.0, 0.0, 0.0, 0.0, 0.05, 0.05, 0.05, 0.05, 0.0, 0.01])
    say("parity", name, "seed", seed)
    w = FTWalkers(Model(seed=seed, **kw))
    o = oracle.OracleChain(seed=seed, **kw)
    nbr = w.tables[0]["lat"]["nnlist"]
    net = EffectiveAction(capacity=16, seed=1)
    net.set_weights(wts)
    net.set_norm(norm)
    w.set_neighbours()
    w.set_proposal(net, nprop=nprop, local_updates=local)
    o.set_proposal(nbr, nprop, wts, norm, local_updates=local)
    w.init_conf()
    o.init_conf()
    w.sweep_start_0b()
    o.sweep_start()
    for s in range(2):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
            cw, co = w.get_conf(), o.conf
            assert np.array_equal(cw, co), "field differs after %s %d (%d spins)" % (p, s, (cw != co).sum())
            assert relmax(w.G(), o.G) < 1e-10, p
    sw, so = w.proposal_stats(), o.proposal_stats()
    say("  stats", sw)
    assert sw == so
    assert 0 < sw["accepted"] < sw["moves"]
    dw, do = w.diag(), o.diag()
    assert dw["accepted"] == do["accepted"]
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert dw["max_wrap_error"] < 1e-8
    ow, oo = w.obs(), o.obs()
    for k in ("sign", "density", "ekin", "dblocc", "s_stag"):
        assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), k
    w.close()
    net.close()


@pytest.mark.parametrize("la,nprop,beta,min_block", [(8, 40, 1.0, 17), (12, 64, 1.0, 33), (20, 64, 0.5, 49)])
def test_block_move_parity_large_blocks(la, nprop, beta, min_block):
    """Accepted blocks of 17-64 sites, so that every depth of the rank-|F| update of G (16, 32,
    48, 64) is taken: a network with all weights zero accepts every inner attempt, and at
    U = 0.5 most blocks pass the exact test.  Field bit-identical with the oracle after every
    pass, G within 1e-10, identical proposal statistics."""
    kw = dict(lattice="square", la=la, lb=la, beta=beta, dtau=0.1, rhub=0.5, kmode="checkerboard", nwrap=5,
              update="delay", nublock=8)
    seed = chain_seed(2)
    from sunhubbardmy_amd.learned import NWEIGHTS
    wts = np.zeros(NWEIGHTS, np.float32)
    norm = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 1.0])
    say("large blocks", la, nprop, beta, "seed", seed)
    w = FTWalkers(Model(seed=seed, **kw))
    o = oracle.OracleChain(seed=seed, **kw)
    net = EffectiveAction(capacity=16, seed=1)
    net.set_weights(wts)
    net.set_norm(norm)
    w.set_neighbours()
    w.set_proposal(net, nprop=nprop, local_updates=False)
    o.set_proposal(w.tables[0]["lat"]["nnlist"], nprop, wts, norm, local_updates=False)
    w.init_conf()
    o.init_conf()
    w.sweep_start_0b()
    o.sweep_start()
    for s in range(2):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
            assert np.array_equal(w.get_conf(), o.conf), (p, s)
            assert relmax(w.G(), o.G) < 1e-10, (p, s)
    sw, so = w.proposal_stats(), o.proposal_stats()
    say("  stats", sw)
    assert sw == so
    assert sw["accepted"] > 0 and sw["spins_accepted"] / sw["accepted"] >= min_block
    w.close()
    net.close()


KW2 = dict(lattice="square", la=2, lb=2, beta=1.0, dtau=0.5, rhub=6.0, kmode="checkerboard", nwrap=1,
           update="delay", nublock=2)


def test_block_moves_detailed_balance():
    """Block moves only, a random effective action worth about one unit per flip: the batch
    samples the exact distribution of the 4-site, 2-slice system, not the effective one."""
    sys.path.insert(0, __file__.rsplit("/", 1)[0])
    from test_oracle_learn import _code, _enumerate
    p, fs, nbr = _enumerate(KW2)
    wts = _rand_weights(11, scale=1.0)
    norm = np.array([0.0, 0.0, 0.0, 0.0, 0.3, 0.3, 0.3, 0.3, 0.0, 0.3])
    e = mlp.forward(wts, norm, fs) * 8
    q = p * np.exp(-(e - e.min()))
    q /= q.sum()
    nch = 32
    say("detailed balance:", nch, "chains")
    w = FTWalkers([Model(seed=chain_seed(c), **KW2) for c in range(nch)])
    net = EffectiveAction(capacity=16, seed=1)
    net.set_weights(wts)
    net.set_norm(norm)
    w.set_neighbours()
    w.set_proposal(net, nprop=6, local_updates=False)
    w.init_conf()
    w.sweep_start_0b()
    for _ in range(50):
        w.sweep(True, False)
    hist = np.zeros(p.size)
    for s in range(1000):
        for part in ("b0", "0b"):
            getattr(w, "sweep_" + part)(True, False)
            for c in range(nch):
                hist[_code(w.get_conf(c))] += 1
        if s % 250 == 0:
            say("  sweep", s)
    hist /= hist.sum()
    tv = 0.5 * np.abs(hist - p).sum()
    tvq = 0.5 * np.abs(hist - q).sum()
    say("  tv", tv, "tv vs effective", tvq)
    assert tv < 0.05 and tvq > 3 * tv
    st = [w.proposal_stats(c) for c in range(nch)]
    assert all(0 < x["accepted"] < x["moves"] for x in st)
    assert max(w.diag(c)["max_wrap_error"] for c in range(nch)) < 1e-10
    w.close()
    net.close()


def test_trained_action_beats_untrained():
    """On a 6x6 lattice: train on configurations of the plain sweep, then compare block
    proposals from the trained network with those of an untrained (zero) one, whose inner chain
    accepts every attempt.  Over 10240 proposals per network the acceptance has a standard
    error of about 0.004."""
    kw = dict(lattice="square", la=6, lb=6, beta=2.0, dtau=0.1, rhub=4.0, kmode="checkerboard", nwrap=10,
              update="delay", nublock=8)
    nch = 64
    w = FTWalkers([Model(seed=chain_seed(c), **kw) for c in range(nch)])
    w.set_neighbours()
    w.init_conf()
    w.sweep_start_0b()
    net = EffectiveAction(capacity=16384, seed=2)
    say("trained vs untrained: sampling")
    for s in range(40):
        w.sweep_b0(True, False)
        if s >= 5:
            net.add_samples(w)
        w.sweep_0b(True, False)
        if s >= 5:
            net.add_samples(w)
    hist = net.train(epochs=60, batch=256, lr=1e-3)
    say("  loss", hist[0], "->", hist[-1])
    assert hist[-1] < hist[0]
    moved, accepted = {}, {}
    for label in ("trained", "untrained"):
        if label == "untrained":
            net.set_weights(np.zeros(mlp.NW, np.float32))
        w.set_proposal(net, nprop=32, local_updates=True)
        w.reset_proposal_stats()
        for _ in range(4):
            w.sweep(True, False)
        st = [w.proposal_stats(c) for c in range(nch)]
        moved[label] = sum(x["spins_accepted"] for x in st) / sum(x["moves"] for x in st)
        accepted[label] = sum(x["accepted"] for x in st) / sum(x["moves"] for x in st)
        say("  ", label, "spins moved per block proposal", moved[label], "acceptance", accepted[label])
    assert accepted["trained"] > accepted["untrained"] + 0.02, accepted
    assert moved["trained"] > 1.1 * moved["untrained"], moved
    w.close()
    net.close()
