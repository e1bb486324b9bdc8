"""GPU parity: the HIP sweep (through the C ABI) against the CPU oracle on identical seeds.

Bar: identical accept/reject sequences (the HS field after every pass is
bit-identical), G within 1e-10 relative max-norm after every stabilization (nwrap), phase within
1e-10, observable bin sums within 1e-9 relative.
"""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model, chain_seed
from sunhubbardmy_amd import _lib

pytestmark = pytest.mark.gpu

BASE = dict(lattice="square", la=4, lb=4, beta=2.0, dtau=0.1, rt=1.0, mu=0.0, rhub=4.0, nu=0.0, nflr=2,
            exact_hs=True, kmode="checkerboard", update="delay", nwrap=10, nublock=4)

CASES = {
    "square4_delay_checkerboard": {},
    "square4_fast_checkerboard": dict(update="fast"),
    "square4_delay_dense": dict(kmode="dense"),
    "square4_submatrix_checkerboard": dict(update="submatrix", nublock=5),
    "square6_su4_general_submatrix_dense": dict(la=6, lb=6, nflr=4, exact_hs=False, rhub=6.0, update="submatrix",
                                                 kmode="dense", nublock=8),
    "honeycomb3_delay_dense": dict(lattice="honeycomb", la=3, lb=3, kmode="dense", nublock=8),
    "honeycomb3_delay_checkerboard_nwrap7": dict(lattice="honeycomb", la=3, lb=3, nwrap=7, beta=1.6),
    "square4_su4_exact_delay": dict(nflr=4, exact_hs=True, rhub=6.0, nublock=3),
    "chain16_delay_checkerboard": dict(lattice="chain", la=16, lb=1, nublock=8),
    "chain10_fast_dense_mu": dict(lattice="chain", la=10, lb=1, kmode="dense", update="fast", mu=0.3),
    "cubic4_delay_checkerboard": dict(lattice="cubic", la=4, lb=4, lc=4, beta=1.0, nublock=8),
    "cubic4_submatrix_dense": dict(lattice="cubic", la=4, lb=4, lc=4, beta=0.8, kmode="dense", update="submatrix",
                                   nublock=8),
    # N = 400 at depth 64: the submatrix kernel stages 3 sites per block (LDS-bound depth) and
    # its flush runs as two passes of 32 (acceptance ~0.8: about 5 full flushes per slice)
    "square20_submatrix64_dense": dict(la=20, lb=20, beta=0.2, kmode="dense", update="submatrix", nublock=64,
                                       nwrap=2),
}


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_parity(kw, nsweeps=2, chain=0, legacy_update=False):
    seed = chain_seed(chain)
    m = Model(seed=seed, **kw)
    say("case", kw, "seed", seed, "legacy_update", legacy_update)
    w = FTWalkers([m])
    if legacy_update:
        a = np.zeros(8, np.int32)
        a[0] = 1
        assert _lib.testing_lib().sh_dbg_op(w._ctx, 14, a.ctypes.data_as(_lib.ip)) == 0
    o = oracle.OracleChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    assert np.array_equal(w.get_conf(), o.conf), "initial HS field differs"
    w.sweep_start_0b()
    o.sweep_start()
    assert relmax(w.G(), o.G) < 1e-10
    cap = 4 * o.nst * nsweeps + 4
    w.set_trace(cap)
    o.set_trace(cap)
    for s in range(nsweeps):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
            say("  sweep", s, p)
            cw, co = w.get_conf(), o.conf
            assert np.array_equal(cw, co), "HS field differs after pass %d %s (%d sites)" % (s, p, (cw != co).sum())
    tw, to = w.trace(), o.trace()
    assert len(tw) == len(to) and len(tw) > 0
    for k in range(len(tw)):
        assert relmax(tw[k], to[k]) < 1e-10, "G differs at stabilization %d: %g" % (k, relmax(tw[k], to[k]))
    dw, do = w.diag(), o.diag()
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert dw["accepted"] == do["accepted"] and dw["tried"] == do["tried"]
    assert dw["max_wrap_error"] < 1e-8
    ow, oo = w.obs(), o.obs()
    assert ow["nobs"] == oo["nobs"]
    for k in ("sign", "density", "ekin", "dblocc", "s_stag"):
        assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), k
    w.close()
    return dw, do


@pytest.mark.parametrize("name", list(CASES))
def test_sweep_parity(name):
    kw = dict(BASE)
    kw.update(CASES[name])
    run_parity(kw)


@pytest.mark.parametrize("name", ["square4_delay_checkerboard", "honeycomb3_delay_dense"])
def test_sweep_parity_l2_delay_kernel(name):
    kw = dict(BASE)
    kw.update(CASES[name])
    run_parity(kw, legacy_update=True)


def test_sweep_parity_metric_shape_short():
    # the bench shape (16x16, k = 32) at a short beta: register-resident delay kernel at K = 32,
    # blocked stabilization at N = 256
    kw = dict(BASE)
    kw.update(la=16, lb=16, beta=1.0, rhub=6.0, nublock=32, nwrap=5)
    run_parity(kw, nsweeps=1)


def test_uneven_batch_matches_single_chains():
    """Chains of different N and lattices in one batch (one launch per kernel for all of them):
    every chain's field and G equal its own oracle run with the same seed."""
    base = dict(beta=1.0, dtau=0.1, rhub=4.0, nflr=2, kmode="checkerboard", update="delay", nwrap=5, nublock=4)
    shapes = [dict(lattice="square", la=4, lb=4), dict(lattice="honeycomb", la=3, lb=3),
              dict(lattice="square", la=6, lb=6), dict(lattice="chain", la=10, lb=1)]
    models = [Model(seed=chain_seed(c), **base, **sh) for c, sh in enumerate(shapes)]
    say("uneven batch", [m.lattice for m in models])
    w = FTWalkers(models)
    oracles = [oracle.OracleChain(seed=chain_seed(c), **base, **sh) for c, sh in enumerate(shapes)]
    w.init_conf()
    for o in oracles:
        o.init_conf()
    w.sweep_start_0b()
    for o in oracles:
        o.sweep_start()
    for p in ("b0", "0b"):
        getattr(w, "sweep_" + p)(True, True)
        for o in oracles:
            getattr(o, "sweep_" + p)(True, True)
    for c, o in enumerate(oracles):
        assert np.array_equal(w.get_conf(c), o.conf), c
        assert relmax(w.G(c), o.G) < 1e-10, c
        ow, oo = w.obs(c), o.obs()
        for k in ("density", "ekin", "dblocc", "s_stag"):
            assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), (c, k)
    w.close()
