"""EffectiveAction: the learned proposal's network (include/sunhip_learn.h).

E_eff(C) = Ns (ymu + ysd g(x(C))) with g a 4 -> 128 -> 128 -> 1 tanh MLP over four global
couplings of the HS field (on-site, nearest neighbours in space, in imaginary time, and
across both).  It is trained on the sweep's own configurations (``add_samples`` after a pass,
``train``) by the library's FP32 kernels, and handed to ``FTWalkers.set_proposal``, which then
runs, after the local update of every slice, a Metropolis chain on E_eff over up to 64 spins of
that slice and accepts the flipped block with the exact fermion determinant ratio.  The
reference has no counterpart (its local updates, src/dqmc_config_u/dqmc_update*.f90, are its
only moves).  Everything runs in libsunhip.so; there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib

HIDDEN = 128
NWEIGHTS = HIDDEN * HIDDEN + 7 * HIDDEN + 1
_fp = ctypes.POINTER(ctypes.c_float)


class LearnError(RuntimeError):
    pass


def weight_views(w):
    """Named views (W1[128,4], b1, W2[128,128], b2, w3, b3) of a flat weight vector."""
    H = HIDDEN
    o = 0
    out = {}
    for name, shape in (("W1", (H, 4)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("w3", (H,)), ("b3", (1,))):
        size = int(np.prod(shape))
        out[name] = w[o:o + size].reshape(shape)
        o += size
    return out


class EffectiveAction:
    def __init__(self, capacity=65536, seed=1, device=0):
        self._net = ctypes.c_void_p()
        rc = _lib.lib().shl_create(ctypes.byref(self._net), int(capacity), int(np.int32(seed)), int(device))
        if rc != 0:
            raise LearnError("shl_create failed (%d)" % rc)
        self.capacity = capacity

    def close(self):
        if getattr(self, "_net", None) and self._net.value:
            _lib.lib().shl_destroy(self._net)
            self._net = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = _lib.lib().shl_last_error(self._net)
            raise LearnError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    # -- parameters --------------------------------------------------------------------------
    def weights(self):
        w = np.zeros(NWEIGHTS, np.float32)
        self._check(_lib.lib().shl_get_weights(self._net, w.ctypes.data_as(_fp)), "get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != NWEIGHTS:
            raise LearnError("expected %d weights" % NWEIGHTS)
        self._check(_lib.lib().shl_set_weights(self._net, w.ctypes.data_as(_fp)), "set_weights")

    def norm(self):
        nm = np.zeros(10)
        self._check(_lib.lib().shl_get_norm(self._net, nm.ctypes.data_as(_lib.dp)), "get_norm")
        return nm

    def set_norm(self, nm):
        nm = np.ascontiguousarray(nm, np.float64)
        self._check(_lib.lib().shl_set_norm(self._net, nm.ctypes.data_as(_lib.dp)), "set_norm")

    # -- samples -----------------------------------------------------------------------------
    def add_samples(self, walkers):
        """One sample per chain: the current field's features and E/Ns (call after a pass)."""
        self._check(_lib.lib().shl_add_samples(self._net, walkers._ctx), "add_samples")

    def push_samples(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 5)
        self._check(_lib.lib().shl_push_samples(self._net, x.ctypes.data_as(_lib.dp), x.shape[0]), "push_samples")

    def num_samples(self):
        return _lib.lib().shl_num_samples(self._net)

    def samples(self):
        n = self.num_samples()
        x = np.zeros((n, 5))
        if n:
            self._check(_lib.lib().shl_get_samples(self._net, x.ctypes.data_as(_lib.dp), 0, n), "get_samples")
        return x

    def clear_samples(self):
        self._check(_lib.lib().shl_clear_samples(self._net), "clear_samples")

    # -- training and evaluation -------------------------------------------------------------
    def train(self, epochs=20, batch=256, lr=1e-3):
        """Adam over shuffled minibatches of the stored samples; per-epoch mean loss."""
        loss = np.zeros(max(epochs, 1))
        self._check(_lib.lib().shl_train(self._net, int(epochs), int(batch), float(lr),
                                         loss.ctypes.data_as(_lib.dp)), "train")
        return loss[:epochs]

    def loss_grad(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 5)
        loss = ctypes.c_double(0.0)
        g = np.zeros(NWEIGHTS, np.float32)
        self._check(_lib.lib().shl_loss_grad(self._net, x.ctypes.data_as(_lib.dp), x.shape[0], ctypes.byref(loss),
                                             g.ctypes.data_as(_fp)), "loss_grad")
        return loss.value, g

    def predict(self, f):
        """E_eff / Ns for features f[n, 4]."""
        f = np.ascontiguousarray(f, np.float64).reshape(-1, 4)
        e = np.zeros(f.shape[0])
        self._check(_lib.lib().shl_predict(self._net, f.ctypes.data_as(_lib.dp), f.shape[0],
                                           e.ctypes.data_as(_lib.dp)), "predict")
        return e
