"""GPU parity of the plaquette field of the finite-temperature sweep (k_plq_slice: the
plaquettes of a slice one after the other, prop and update, k_plq_prop: every plaquette on the
rows or columns of a B product) with the oracle through the C ABI: the plaquette field (and the
on-site and bond fields where present) identical after every pass, G within 1e-10 of the oracle
at every stabilization, the phase, the acceptance sums of every field and the observable bins
equal; the field's get / set and acceptance counters; a batch refuses chains with and without
the field."""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model, chain_seed
from sunhubbardmy_amd.walkers import SweepError

pytestmark = pytest.mark.gpu

CASES = {
    "square4_plq_only_checkerboard": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=0.0, rhub_plq=1.5, alpha=0.3,
                                          kmode="checkerboard", nwrap=5, update="delay", nublock=4),
    "square4_hubbard_plq_delay": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=3.0, rhub_plq=1.5, alpha=0.3,
                                      kmode="checkerboard", nwrap=5, update="delay", nublock=4),
    "square4_su4_plq_theta_dense": dict(lattice="square", la=4, lb=4, beta=0.6, rhub=2.0, nflr=4, exact_hs=False,
                                        rhub_plq=1.0, alpha=0.4, theta=0.5, kmode="dense", nwrap=3, update="fast"),
    "honeycomb3_plq_submatrix": dict(lattice="honeycomb", la=3, lb=3, beta=0.8, rhub=3.0, rhub_plq=1.2, alpha=0.3,
                                     theta=0.25, kmode="dense", nwrap=4, update="submatrix", nublock=8),
    "honeycomb3_tuv_plq_delay": dict(lattice="honeycomb", la=3, lb=3, beta=0.6, rhub=3.0, rv=1.0, rhub_plq=1.0,
                                     alpha=0.3, kmode="checkerboard", nwrap=3, update="delay", nublock=4),
    "square6_plq_nwrap1": dict(lattice="square", la=6, lb=6, beta=0.4, rhub=4.0, rhub_plq=2.0, alpha=0.2,
                               kmode="checkerboard", nwrap=1, update="delay", nublock=8),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_plq_parity(name):
    kw = CASES[name]
    seed = chain_seed(9)
    say("case", name, "seed", seed)
    w = FTWalkers([Model(seed=seed, **kw)])
    o = oracle.OracleChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    assert np.array_equal(w.get_confp(), o.confp), "initial plaquette field differs"
    if kw["rhub"]:
        assert np.array_equal(w.get_conf(), o.conf)
    w.sweep_start_0b()
    o.sweep_start()
    assert rel(w.G(), o.G) < 1e-10
    cap = 4 * o.nst * 2 + 4
    w.set_trace(cap)
    o.set_trace(cap)
    for s in range(2):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
            say("  sweep", s, p)
            assert np.array_equal(w.get_confp(), o.confp), "plaquette field differs after %d %s" % (s, p)
            if kw["rhub"]:
                assert np.array_equal(w.get_conf(), o.conf), "on-site field differs after %d %s" % (s, p)
            if kw.get("rv"):
                assert np.array_equal(w.get_confv(), o.confv), "bond field differs after %d %s" % (s, p)
    tw, to = w.trace(), o.trace()
    assert len(tw) == len(to) > 0
    for k in range(len(tw)):
        assert rel(tw[k], to[k]) < 1e-10, "G differs at stabilization %d: %g" % (k, rel(tw[k], to[k]))
    dw, do = w.diag(), o.diag()
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert dw["max_wrap_error"] < 1e-8 and do["max_wrap_error"] < 1e-8
    assert np.array_equal(w.pacc(), o.pacc()) and 0 < w.pacc()[0] < w.pacc()[1]
    if kw.get("rv"):
        assert np.array_equal(w.vacc(), o.vacc())
    ow, oo = w.obs(), o.obs()
    assert ow["nobs"] == oo["nobs"]
    for k in ("sign", "density", "ekin", "dblocc", "s_stag"):
        assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), k
    w.close()


def test_plq_field_io_and_batch_rules():
    kw = CASES["square4_hubbard_plq_delay"]
    w = FTWalkers([Model(seed=chain_seed(c), **kw) for c in range(2)])
    w.init_conf()
    c = w.get_confp(1)
    assert c.shape == (10, 16) and set(np.unique(c)) == {1, 2, 3, 4}
    c2 = (c % 4) + 1
    w.set_confp(c2, 1)
    assert np.array_equal(w.get_confp(1), c2)
    with pytest.raises(SweepError):
        w.set_confp(np.zeros_like(c2), 0)           # values 1..4 only
    assert np.array_equal(w.pacc(0), [0.0, 0.0])
    w.close()
    plain = dict(kw, rhub_plq=0.0)
    with pytest.raises(SweepError):
        FTWalkers([Model(seed=1, **kw), Model(seed=2, **plain)])


def test_plq_resume_bit_exact(tmp_path):
    """save_state after one sweep, then one more sweep, equals a fresh batch restored from the
    file: the plaquette field travels with the on-site field and the generator states"""
    from sunhubbardmy_amd import confio
    kw = dict(CASES["square4_hubbard_plq_delay"])
    ms = [Model(seed=chain_seed(c), **kw) for c in range(2)]
    a = FTWalkers(ms)
    a.init_conf()
    a.sweep_start_0b()
    a.sweep(lupdate=True, lmeasure=False)
    path = str(tmp_path / "state.npz")
    confio.save_state(path, a)
    a.sweep(lupdate=True, lmeasure=False)
    b = FTWalkers(ms)
    confio.load_state(path, b)
    b.sweep_start_0b()
    b.sweep(lupdate=True, lmeasure=False)
    for c in range(2):
        assert np.array_equal(a.get_confp(c), b.get_confp(c)) and np.array_equal(a.get_conf(c), b.get_conf(c))
        assert rel(a.G(c), b.G(c)) < 1e-10
    a.close()
    b.close()
