"""GPU parity of the correlation bins of the equal-time measurement (k_eqt) with the oracle, through
the C ABI: the same seeds give the same HS field, and the bins summed over two measured sweeps agree
to 1e-9 relative for the square (rectangular), honeycomb, chain and cubic lattices."""
import sys

import numpy as np
import pytest

import oracle
from sunhubbardmy_amd import FTWalkers, Model, binio, chain_seed

pytestmark = pytest.mark.gpu

THETA = 0.3 * np.pi
CASES = {
    "square4x6_delay": dict(lattice="square", la=4, lb=6, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=5,
                            update="delay", nublock=4),
    "honeycomb3_dense": dict(lattice="honeycomb", la=3, lb=3, beta=1.0, rhub=3.0, kmode="dense", nwrap=5,
                             update="fast"),
    "chain10_delay": dict(lattice="chain", la=10, lb=1, beta=1.0, rhub=4.0, kmode="checkerboard", nwrap=5,
                          update="delay", nublock=4),
    "cubic4x4x2_submatrix": dict(lattice="cubic", la=4, lb=4, lc=2, beta=0.6, rhub=4.0, kmode="checkerboard",
                                 nwrap=3, update="submatrix", nublock=8),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", CASES)
def test_eqt_parity(name):
    kw = CASES[name]
    seed = chain_seed(6)
    say("case", name, "seed", seed)
    w = FTWalkers([Model(seed=seed, **kw)])
    o = oracle.OracleChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    w.set_eqt(THETA)
    o.set_eqt(THETA)
    w.sweep_start_0b()
    o.sweep_start()
    for _ in range(2):
        for p in ("b0", "0b"):
            getattr(w, "sweep_" + p)(True, True)
            getattr(o, "sweep_" + p)(True, True)
    assert np.array_equal(w.get_conf(), o.conf)
    bw, bo = w.eqt(), o.eqt()
    uw, uo = binio.unpack_eqt(bw, o.lq), binio.unpack_eqt(bo, o.lq)
    for k in binio.KEYS22 + binio.KEYS1 + binio.SCALARS:
        a, b = np.asarray(uw[k]), np.asarray(uo[k])
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1.0)
        say("  ", k, err)
        assert err < 1e-9, k
    ow, oo = w.obs(), o.obs()
    assert ow["nobs"] == oo["nobs"]
    w.reset_obs()
    assert not np.any(w.eqt())
    w.close()


@pytest.mark.parametrize("name", ["honeycomb2_tv_dense", "square4_hubbard_cb"])
def test_eqt_parity_projective(name):
    from sunhubbardmy_amd.proj import ProjModel, ProjWalkers
    kw = {"honeycomb2_tv_dense": dict(lattice="honeycomb", la=2, lb=2, beta=1.0, rv=1.2, rndness=0.2, nflr=1,
                                      kmode="dense", nwrap=5, nustock=8, mid_len=2),
          "square4_hubbard_cb": dict(lattice="square", la=4, lb=4, beta=1.0, rv=0.0, rhub=4.0, nflr=2, rndness=0.2,
                                     kmode="checkerboard", nwrap=4, nustock=6, mid_len=2)}[name]
    seed = chain_seed(3)
    say("case", name, "seed", seed)
    w = ProjWalkers([ProjModel(seed=seed, **kw)])
    o = oracle.OracleProjChain(seed=seed, **kw)
    w.init_conf()
    o.init_conf()
    w.set_trial()
    o.set_trial()
    o.set_proj(w.projs[0])
    w.set_eqt(THETA)
    o.set_eqt(THETA)
    w.sweep_start_0b()
    o.sweep_start()
    for _ in range(2):
        w.sweep_b0(True, True)
        o.sweep_b0(True, True)
        w.sweep_0b(True, True)
        o.sweep_0b(True, True)
    assert np.array_equal(w.get_conf(), o.conf)
    uw, uo = binio.unpack_eqt(w.eqt(), o.lq), binio.unpack_eqt(o.eqt(), o.lq)
    assert o.obs()["nobs"] > 0
    for k in binio.KEYS22 + binio.KEYS1 + binio.SCALARS:
        a, b = np.asarray(uw[k]), np.asarray(uo[k])
        err = np.abs(a - b).max() / max(np.abs(b).max(), 1.0)
        say("  ", k, err)
        assert err < 1e-9, k
    w.close()
