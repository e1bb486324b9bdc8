"""GPU parity of the finite-temperature t-V bond field (k_vprop, the B products with the field,
and the bond update: the delay form k_vupdate_delay in delay builds, the submatrix form
k_vupdate_submatrix in submatrix builds, the fast form k_vupdate otherwise) with the oracle's fast
form, through the C ABI: the two algorithms take the same
decisions, so the on-site and bond fields are identical after every pass, G agrees to 1e-10 at
every stabilization, and the phase, the acceptance sums and the observable bins agree."""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model, chain_seed

pytestmark = pytest.mark.gpu

CASES = {
    "square4_tv_checkerboard": dict(lattice="square", la=4, lb=4, beta=1.0, rhub=0.0, rv=1.5, nflr=1,
                                    kmode="checkerboard", nwrap=5, update="delay", nublock=4),
    "chain16_tv_dense": dict(lattice="chain", la=16, lb=1, beta=1.0, rhub=0.0, rv=2.0, nflr=1, mu=0.2,
                             kmode="dense", nwrap=4, update="fast"),
    "honeycomb3_tuv_delay": dict(lattice="honeycomb", la=3, lb=3, beta=1.0, rhub=3.0, rv=1.0, nflr=2,
                                 kmode="checkerboard", nwrap=5, update="delay", nublock=4),
    "square4_tuv_submatrix_dense": dict(lattice="square", la=4, lb=4, beta=0.8, rhub=2.0, rv=1.0, nflr=2,
                                        kmode="dense", nwrap=4, update="submatrix", nublock=4),
    "square6_tv_delay16": dict(lattice="square", la=6, lb=6, beta=1.0, rhub=0.0, rv=1.5, nflr=1,
                               kmode="checkerboard", nwrap=5, update="delay", nublock=16),
    "honeycomb4_tuv_delay32_dense": dict(lattice="honeycomb", la=4, lb=4, beta=0.6, rhub=3.0, rv=1.2, nflr=2,
                                         kmode="dense", nwrap=3, update="delay", nublock=32),
    "chain16_tv_delay6": dict(lattice="chain", la=16, lb=1, beta=1.0, rhub=0.0, rv=2.0, nflr=1,
                              kmode="checkerboard", nwrap=5, update="delay", nublock=6),
    # the submatrix form k_vupdate_submatrix: 2x2 bordering of Gamma^-1, flush every K entries
    # (K = nublock rounded down to even) and after each family's last bond
    "square6_tv_submatrix16": dict(lattice="square", la=6, lb=6, beta=1.0, rhub=0.0, rv=1.5, nflr=1,
                                   kmode="checkerboard", nwrap=5, update="submatrix", nublock=16),
    "honeycomb4_tuv_submatrix32_dense": dict(lattice="honeycomb", la=4, lb=4, beta=0.6, rhub=3.0, rv=1.2, nflr=2,
                                             kmode="dense", nwrap=3, update="submatrix", nublock=32),
    "chain16_tv_submatrix5": dict(lattice="chain", la=16, lb=1, beta=1.0, rhub=0.0, rv=2.0, nflr=1,
                                  kmode="checkerboard", nwrap=5, update="submatrix", nublock=5),
    "square8_tv_submatrix64": dict(lattice="square", la=8, lb=8, beta=0.5, rhub=0.0, rv=1.5, nflr=1,
                                   kmode="checkerboard", nwrap=5, update="submatrix", nublock=64),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_tv_parity(name):
    kw = CASES[name]
    seed = chain_seed(7)
    say("case", name, "seed", seed)
    w = FTWalkers([Model(seed=seed, **kw)])
    o = oracle.OracleChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    assert np.array_equal(w.get_confv(), o.confv), "initial bond field differs"
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
            assert np.array_equal(w.get_confv(), o.confv), "bond field differs after %d %s" % (s, p)
            if kw["rhub"]:
                assert np.array_equal(w.get_conf(), o.conf), "on-site field differs after %d %s" % (s, p)
    tw, to = w.trace(), o.trace()
    assert len(tw) == len(to) > 0
    for k in range(len(tw)):
        assert rel(tw[k], to[k]) < 1e-10, "G differs at stabilization %d: %g" % (k, rel(tw[k], to[k]))
    dw, do = w.diag(), o.diag()
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert dw["max_wrap_error"] < 1e-8 and do["max_wrap_error"] < 1e-8
    assert np.array_equal(w.vacc(), o.vacc()) and 0 < w.vacc()[0] < w.vacc()[1]
    ow, oo = w.obs(), o.obs()
    assert ow["nobs"] == oo["nobs"]
    for k in ("sign", "density", "ekin", "dblocc", "s_stag"):
        assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), k
    w.close()


def test_tv_resume_bit_exact(tmp_path):
    """save_state after one sweep, then one more sweep, equals a fresh batch restored from the file"""
    from sunhubbardmy_amd import confio
    kw = dict(CASES["honeycomb3_tuv_delay"])
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
        assert np.array_equal(a.get_confv(c), b.get_confv(c)) and np.array_equal(a.get_conf(c), b.get_conf(c))
        assert rel(a.G(c), b.G(c)) < 1e-10
    a.close()
    b.close()
