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
