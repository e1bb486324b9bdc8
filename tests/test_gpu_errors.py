"""Error paths of the library, forced from the host so that every kernel still runs to its end.

  non-finite G       NaN written into the stored right UDV chain (D of every chain point)
                     makes the Green's function of the next stabilization non-finite: the pass
                     completes (no fault, no abort), sh_get_diag returns SH_E_NONFINITE, and a
                     fresh sweep_start_0b plus sh_reset_diag recover the chain
  LU breakdown       an exactly singular matrix in the LU of green_equaltime returns the first
                     zero pivot + 1 (LAPACK's info, the reference's `stop` condition in
                     dqmc_ft_sweep_b0.f90:188-194)
  call order         a pass before sweep_start_0b is SH_E_STATE
"""
import numpy as np
import pytest

from sunhubbardmy_amd import FTWalkers, Model, SweepError, chain_seed
from sunhubbardmy_amd import _lib

pytestmark = pytest.mark.gpu

KW = dict(lattice="square", la=4, lb=4, beta=1.0, dtau=0.1, rhub=4.0, kmode="checkerboard", update="delay",
          nublock=4, nwrap=5)


def test_nonfinite_green_function_is_reported_and_recoverable():
    w = FTWalkers([Model(seed=chain_seed(0), **KW)])
    L = _lib.testing_lib()
    w.init_conf()
    w.sweep_start_0b()
    n, nst = w.tables[0]["ndim"], w.tables[0]["nst"]
    nan = np.full(n, np.nan)
    for pt in range(nst + 1):
        assert L.sh_dbg_put(w._ctx, 0, 100, pt, nan.ctypes.data_as(_lib.dp)) == 0
    print("nonfinite: NaN chain points written", flush=True)
    w.sweep_b0(True, True)                 # runs to its end
    with pytest.raises(SweepError) as e:
        w.diag()
    assert "(-4)" in str(e.value)
    assert not np.all(np.isfinite(w.G()))
    # recovery: a fresh start from the HS field and cleared diagnostics
    w.reset_diag()
    w.sweep_start_0b()
    w.sweep(True, True)
    d = w.diag()
    assert np.all(np.isfinite(w.G())) and d["max_wrap_error"] < 1e-8
    w.close()


@pytest.mark.parametrize("zero_col", [0, 5, 15])
def test_lu_breakdown_reports_the_first_zero_pivot(zero_col):
    m = Model(seed=chain_seed(0), **KW)
    w = FTWalkers([m])
    L = _lib.testing_lib()
    n = w.tables[0]["ndim"]
    rng = np.random.default_rng(zero_col)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 2 * np.eye(n)
    M[:, zero_col] = 0.0
    X = rng.standard_normal((n, n)) + 0j
    for sel, a in ((4, M), (5, X)):
        a = np.asfortranarray(a, dtype=np.complex128)
        assert L.sh_dbg_put(w._ctx, 0, sel, 0, a.ctypes.data_as(_lib.dp)) == 0
    a = np.zeros(8, np.int32)
    assert L.sh_dbg_op(w._ctx, 5, a.ctypes.data_as(_lib.ip)) == 0
    sc = np.zeros(32)
    assert L.sh_dbg_get(w._ctx, 0, 104, 0, sc.ctypes.data_as(_lib.dp)) == 0
    # partial pivoting meets the zero column at step zero_col: info = zero_col + 1
    assert sc[14] == zero_col + 1
    w.close()


def test_pass_before_start_is_a_state_error():
    w = FTWalkers([Model(seed=chain_seed(0), **KW)])
    w.init_conf()
    with pytest.raises(SweepError) as e:
        w.sweep_b0(True, True)
    assert "(-3)" in str(e.value)
    w.close()
