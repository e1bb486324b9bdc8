"""Graph-captured passes and the autotuner.

Graphs: a batch run with every pass replayed from a captured hipGraph against the same batch
run with direct launches, pass by pass: HS fields, G and the observable sums must be bit-identical
(the same kernels on the same data), across several replays of each captured graph with eager
work and synchronisations between them, and across a setter that retires the graphs.

Autotuner: it leaves the chains' state bit for bit as it was, picks one of its candidates, and a
sweep after it equals a sweep of a batch created with that depth (the depth changes only the
rounding of the delayed update, and identical depths give identical results).
"""
import numpy as np
import pytest

from sunhubbardmy_amd import FTWalkers, Model, chain_seed

pytestmark = pytest.mark.gpu

CASES = {
    "square6_delay_checkerboard": dict(lattice="square", la=6, lb=6, beta=2.0, rhub=4.0, kmode="checkerboard",
                                       update="delay", nublock=8, nwrap=10),
    "square8_su4_submatrix_fft": dict(lattice="square", la=8, lb=8, beta=1.0, rhub=6.0, nflr=4, exact_hs=False,
                                      kmode="fft", update="submatrix", nublock=16, nwrap=5),
    "honeycomb3_delay_dense_tv": dict(lattice="honeycomb", la=3, lb=3, beta=1.0, rhub=4.0, rv=1.0, kmode="dense",
                                      update="delay", nublock=4, nwrap=5),
}


def _batch(kw, nch=3):
    return FTWalkers([Model(seed=chain_seed(c), **kw) for c in range(nch)])


def _same(a, b, nch):
    for c in range(nch):
        assert np.array_equal(a.get_conf(c), b.get_conf(c)), c
        assert np.array_equal(a.G(c), b.G(c)), c
        oa, ob = a.obs(c), b.obs(c)
        for k in ("nobs", "sign", "density", "ekin", "dblocc", "s_stag"):
            assert oa[k] == ob[k], (c, k)


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replay_matches_direct_launches(name):
    kw = CASES[name]
    eager, graph = _batch(kw), _batch(kw)
    graph.set_graphs(True)
    for w in (eager, graph):
        w.init_conf()
        w.sweep_start_0b()
    # each kind of pass replayed three times, with host readbacks between the replays
    plan = [("b0", True, True), ("0b", True, True), ("b0", True, False), ("0b", True, True), ("b0", True, True),
            ("0b", False, True), ("b0", True, True), ("0b", True, True), ("0b", True, True)]
    for i, (p, lupd, lmeas) in enumerate(plan):
        for w in (eager, graph):
            getattr(w, "sweep_" + p)(lupd, lmeas)
        print("pass", i, p, lupd, lmeas, flush=True)
        _same(eager, graph, 3)
        if i == 4:
            # a setter in the middle retires the captured graphs; the next passes capture again
            for w in (eager, graph):
                w.set_trace(8)
    eager.close()
    graph.close()


def test_graph_replay_with_block_proposals():
    from sunhubbardmy_amd.learned import NWEIGHTS, EffectiveAction
    kw = dict(lattice="square", la=4, lb=4, beta=1.0, rhub=4.0, kmode="checkerboard", update="delay", nublock=4,
              nwrap=5)
    wts = (np.random.default_rng(13).standard_normal(NWEIGHTS) * np.sqrt(2.0 / 132)).astype(np.float32)
    norm = np.array([0.0, 0.0, 0.0, 0.0, 0.05, 0.05, 0.05, 0.05, 0.0, 0.01])
    net = EffectiveAction(capacity=16, seed=1)
    net.set_weights(wts)
    net.set_norm(norm)
    eager, graph = _batch(kw, 2), _batch(kw, 2)
    graph.set_graphs(True)
    for w in (eager, graph):
        w.set_neighbours()
        w.set_proposal(net, nprop=16)
        w.init_conf()
        w.sweep_start_0b()
    for s in range(3):
        for w in (eager, graph):
            w.sweep(True, True)
        _same(eager, graph, 2)
        for c in range(2):
            assert eager.proposal_stats(c) == graph.proposal_stats(c)
    eager.close()
    graph.close()
    net.close()


@pytest.mark.parametrize("kw", [
    dict(lattice="square", la=8, lb=8, beta=1.0, rhub=4.0, kmode="checkerboard", update="delay", nublock=32, nwrap=5),
    dict(lattice="square", la=8, lb=8, beta=1.0, rhub=6.0, nflr=4, exact_hs=False, kmode="fft", update="submatrix",
         nublock=32, nwrap=5),
    dict(lattice="honeycomb", la=12, lb=12, beta=0.5, rhub=4.0, kmode="dense", update="delay", nublock=32, nwrap=5),
])
def test_autotune_keeps_state_and_matches_a_batch_created_with_its_choice(kw):
    nch = 4
    w = _batch(kw, nch)
    w.init_conf()
    w.sweep_start_0b()
    g0 = [w.G(c) for c in range(nch)]
    f0 = [w.get_conf(c) for c in range(nch)]
    d0 = [w.diag(c) for c in range(nch)]
    tuned = w.autotune()
    print("autotuned", kw["lattice"], kw["update"], kw["kmode"], tuned, flush=True)
    k = tuned["nublock"]
    cands = {4, 8, 16, 32, 64} if kw["update"] == "delay" else {8, 16, 24, 32, 48, 64}
    assert k in cands
    if kw["kmode"] == "fft":
        assert tuned["fft_vectors_per_workgroup"] in (2, 4, 8, 16)
    for c in range(nch):
        assert np.array_equal(w.G(c), g0[c])
        assert np.array_equal(w.get_conf(c), f0[c])
        assert w.diag(c) == d0[c]
    # the same chains created with nublock = k reach the same state at sweep_start_0b (which
    # does not depend on the depth); one sweep later both batches are bit-identical
    kw2 = dict(kw)
    kw2["nublock"] = k
    ref = _batch(kw2, nch)
    ref.init_conf()
    ref.sweep_start_0b()
    for x in (w, ref):
        x.sweep(True, True)
    _same(w, ref, nch)
    for c in range(nch):
        assert w.diag(c)["max_wrap_error"] < 1e-8
    w.close()
    ref.close()
