"""GPU parity of the projective (submatrix-LR) sweep through the C ABI of include/sunhip_proj.h
against the CPU oracle on identical seeds: the t-V bond field, the Hubbard on-site field and
both together.

Bar: the bond field after every pass is bit-identical (identical accept/reject sequences), G
after every stabilization within 1e-10 relative max-norm, the phase within 1e-10, the
acceptance counts equal and the observable bin sums within 1e-9 relative.  The trial state is
built on the host from the chain's own draws (ProjWalkers.set_trial) and handed to the oracle,
which draws the same numbers, so both start from the same Slater determinant.
"""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import chain_seed
from sunhubbardmy_amd.proj import ProjModel, ProjWalkers

pytestmark = pytest.mark.gpu

BASE = dict(lattice="honeycomb", la=2, lb=2, beta=1.0, dtau=0.1, rt=1.0, mu=0.0, rv=1.2, nu=0.0, rndness=0.2,
            nflr=1, kmode="dense", nwrap=5, nustock=8, mid_len=2)

CASES = {
    "honeycomb2_dense": {},
    "honeycomb3_checkerboard_stock6": dict(la=3, lb=3, kmode="checkerboard", nustock=6),
    "square4_checkerboard_nwrap3": dict(lattice="square", la=4, lb=4, kmode="checkerboard", nwrap=3, nustock=4),
    "honeycomb3_dense_stock64_v2": dict(la=3, lb=3, nustock=64, rv=2.0, beta=1.2, nwrap=4),
    # the Hubbard on-site field (src/dqmc_config_u): alone, with 4 components, and with bonds
    "square4_hubbard_cb": dict(lattice="square", la=4, lb=4, rv=0.0, rhub=4.0, nflr=2, kmode="checkerboard",
                               nustock=6, nwrap=4),
    "square4_su4_dense": dict(lattice="square", la=4, lb=4, rv=0.0, rhub=3.0, nflr=4, kmode="dense", nustock=8),
    "honeycomb3_tuv": dict(la=3, lb=3, rv=1.0, rhub=3.0, nflr=2, nustock=8, kmode="checkerboard"),
    "chain16_tuv_cb": dict(lattice="chain", la=16, lb=1, rv=1.0, rhub=2.0, nflr=2, nustock=8, kmode="checkerboard"),
}


def relmax(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_parity(kw, nsweeps=2, chain=1, gpu_kw=None):
    seed = chain_seed(chain)
    say("case", kw, gpu_kw or "", "seed", seed)
    w = ProjWalkers(ProjModel(seed=seed, **kw, **(gpu_kw or {})))
    o = oracle.OracleProjChain(seed=seed, wrap_check=True, **kw)
    w.init_conf()
    o.init_conf()
    assert np.array_equal(w.get_conf(), o.conf), "initial bond field differs"
    if kw.get("rhub", 0.0) != 0.0:
        assert np.array_equal(w.get_uconf(), o.uconf), "initial on-site field differs"
    w.set_trial()
    o.set_trial()                      # the oracle's stream advances by the same draws
    o.set_proj(w.projs[0])
    w.sweep_start_0b()
    o.sweep_start()
    assert relmax(w.G(), o.G) < 1e-10
    dw, do = w.diag(), o.diag()
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert abs(dw["logweightf"] - do["logweightf"]) < 1e-9 * max(1.0, abs(do["logweightf"]))
    cap = 4 * o.nst * nsweeps + 4
    w.set_trace(cap)
    o.set_trace(cap)
    for s in range(nsweeps):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
            say("  sweep", s, p)
            cw, co = w.get_conf(), o.conf
            assert np.array_equal(cw, co), "bond field differs after pass %d %s (%d bonds)" % (s, p, (cw != co).sum())
            if kw.get("rhub", 0.0) != 0.0:
                uw, uo = w.get_uconf(), o.uconf
                assert np.array_equal(uw, uo), "on-site field differs after pass %d %s (%d)" % (s, p, (uw != uo).sum())
    tw, to = w.trace(), o.trace()
    assert len(tw) == len(to) and len(tw) > 0
    for k in range(len(tw)):
        assert relmax(tw[k], to[k]) < 1e-10, "G differs at stabilization %d: %g" % (k, relmax(tw[k], to[k]))
    assert relmax(w.G(), o.G) < 1e-10
    dw, do = w.diag(), o.diag()
    assert abs(dw["phase"] - do["phase"]) < 1e-10
    assert abs(dw["logweightf"] - do["logweightf"]) < 1e-9 * max(1.0, abs(do["logweightf"]))
    assert dw["accepted"] == do["accepted"] and dw["tried"] == do["tried"]
    assert 0 < dw["accepted"] < dw["tried"]
    assert dw["max_wrap_error"] < 1e-8 and dw["nchecks"] == do["nchecks"]
    ow, oo = w.obs(), o.obs()
    assert ow["nobs"] == oo["nobs"] and ow["nobs"] > 0
    for k in ("sign", "density", "ekin", "dblocc", "s_stag"):
        assert abs(ow[k] - oo[k]) <= 1e-9 * max(1.0, abs(oo[k])), k
    w.close()


@pytest.mark.parametrize("variant,gblock", [("fast", 0), ("delay", 4), ("delay", 32), ("submatrix_g", 8),
                                             ("submatrix_g", 64), ("submatrix_g", 1), ("submatrix_g", 5)])
@pytest.mark.parametrize("name", ["square4_hubbard_cb", "square4_su4_dense", "honeycomb2_dense",
                                  "square4_checkerboard_nwrap3"])
def test_proj_update_variants(name, variant, gblock):
    """The fast, delay and submatrix-G forms of the on-site and the bond update
    (src/dqmc_config_u/ and src/dqmc_config_v/ dqmc_proj_update.f90, dqmc_proj_update_delay.f90,
    dqmc_proj_update_submatrix_G.f90) against the oracle's submatrix-LR form on the same draws:
    the four algorithms take the same decisions, so the fields stay bit-identical and G agrees
    to 1e-10 (rounding of different operation orders)."""
    kw = dict(BASE)
    kw.update(CASES[name])
    run_parity(kw, gpu_kw=dict(update=variant, gblock=gblock))


@pytest.mark.parametrize("name", list(CASES))
def test_proj_parity(name):
    kw = dict(BASE)
    kw.update(CASES[name])
    run_parity(kw)


def test_proj_batch_matches_single_chains():
    """Chains of one batch advance exactly as they would alone."""
    kw = dict(BASE)
    seeds = [chain_seed(c) for c in range(3)]
    w = ProjWalkers([ProjModel(seed=s, **kw) for s in seeds])
    w.init_conf()
    w.set_trial()
    w.sweep_start_0b()
    w.sweep()
    for c, s in enumerate(seeds):
        o = oracle.OracleProjChain(seed=s, wrap_check=True, **kw)
        o.init_conf()
        o.set_trial()
        o.set_proj(w.projs[c])
        o.sweep_start()
        o.sweep()
        assert np.array_equal(w.get_conf(c), o.conf)
        assert relmax(w.G(c), o.G) < 1e-10
    w.close()


def test_proj_full_size_properties():
    """The L = 18 honeycomb workload (N = 648, ne = 324, nustock 64) at its projection length theta = 20:
    UL UR stays invertible, G is a projector of rank N - ne (G^2 = G, tr G = N - ne), the wrap
    error stays small and the acceptance is a proper fraction."""
    m = ProjModel(lattice="honeycomb", la=18, lb=18, beta=20.0, rv=1.2, rndness=1e-4, kmode="dense", nwrap=10,
                  nustock=0, seed=chain_seed(0))
    w = ProjWalkers(m)
    assert (w.ndim, w.ne) == (648, 324)
    w.init_conf()
    w.set_trial()
    w.sweep_start_0b()
    w.sweep()
    g = w.G()
    assert abs(np.trace(g).real - (w.ndim - w.ne)) < 1e-8
    assert np.abs(g @ g - g).max() < 1e-8
    d = w.diag()
    assert d["max_wrap_error"] < 1e-6
    assert 0 < d["accepted"] < d["tried"]
    w.close()
