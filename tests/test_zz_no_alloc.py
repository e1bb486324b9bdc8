"""shl_loss_grad_noalloc, the entry point that allocates nothing, against shl_loss_grad.

The caller owns device buffers for the samples, the loss and the gradient (allocated here with
the HIP runtime the library itself uses); the entry point keeps its partial sums in the
library's per-GPU scratch array.  Two calls in a row, the second with fewer samples, must give
exactly what the allocating entry point gives on the same samples (same kernels, same
partition into workgroups, same summation order): the second call must not see the first
call's partial sums.
"""
import ctypes

import numpy as np
import pytest

from sunhubbardmy_amd import _lib
from sunhubbardmy_amd.learned import NWEIGHTS, EffectiveAction

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.lib()                                # the runtime libsunhip.so is linked against
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _alloc(h, nbytes):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def test_noalloc_loss_grad_matches_the_allocating_entry_point_twice():
    h = _hip()
    L = _lib.lib()
    nmax = 700
    net = EffectiveAction(capacity=nmax, seed=3)
    rng = np.random.default_rng(5)
    net.set_weights((rng.standard_normal(NWEIGHTS) * 0.1).astype(np.float32))
    net.set_norm(np.array([0.1, -0.2, 0.3, 0.0, 1.5, 0.7, 1.1, 2.0, -0.4, 0.8]))
    dx = _alloc(h, nmax * 5 * 8)
    dl = _alloc(h, 8)
    dg = _alloc(h, NWEIGHTS * 4)
    try:
        for n in (nmax, 130):
            print("noalloc loss_grad, n =", n, flush=True)
            x = np.ascontiguousarray(rng.standard_normal((n, 5)))
            ref_loss, ref_grad = net.loss_grad(x)
            assert h.hipMemcpy(dx, x.ctypes.data, x.nbytes, H2D) == 0
            rc = L.shl_loss_grad_noalloc(net._net, dx, n, dl, dg)
            assert rc == 0, L.shl_last_error(net._net)
            loss = np.zeros(1)
            grad = np.zeros(NWEIGHTS, np.float32)
            assert h.hipMemcpy(loss.ctypes.data, dl, 8, D2H) == 0
            assert h.hipMemcpy(grad.ctypes.data, dg, grad.nbytes, D2H) == 0
            assert np.isfinite(ref_loss) and np.abs(ref_grad).max() > 0
            assert loss[0] == ref_loss
            assert np.array_equal(grad, ref_grad)
        # sizes beyond the maximum and empty inputs are refused before any launch
        assert L.shl_loss_grad_noalloc(net._net, dx, 0, dl, dg) == -1
        assert L.shl_loss_grad_noalloc(net._net, dx, (1 << 24) + 1, dl, dg) == -1
        assert b"16777216" in L.shl_last_error(net._net)
    finally:
        for p in (dx, dl, dg):
            h.hipFree(p)
        net.close()
