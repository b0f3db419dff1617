"""ctypes binding of the U-Net entry points (ls_net_*, include/lsgpu.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import unet_spec as S
from .lsgpu import load_library

_sig_done = False


def _lib():
    global _sig_done
    lib = load_library()
    if not _sig_done:
        vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
        fp = C.POINTER(C.c_float)
        sig = {
            "ls_net_num_weights": (sz, []),
            "ls_net_create": (i, [C.POINTER(vp), i, i, i, i, i]),
            "ls_net_destroy": (None, [vp]),
            "ls_net_last_error": (C.c_char_p, [vp]),
            "ls_net_set_weights": (i, [vp, fp, sz]),
            "ls_net_get_weights": (i, [vp, fp, sz]),
            "ls_net_forward": (i, [vp, fp, fp]),
            "ls_net_backward": (i, [vp, fp, fp]),
            "ls_net_train_step": (i, [vp, fp, fp, fp, C.c_float, C.POINTER(C.c_double)]),
            "ls_net_time_forward": (i, [vp, i, C.POINTER(C.c_double)]),
            "ls_net_set_conv": (i, [vp, i]),
            "ls_net_step": (i, [vp, fp, C.c_float]),
            "ls_net_store_alloc": (i, [vp, i, i]),
            "ls_net_store_operator": (i, [vp, i, fp]),
            "ls_net_store_sample": (i, [vp, i, i, fp, fp, C.c_double]),
            "ls_net_energy_step": (i, [vp, C.POINTER(C.c_int), C.c_float,
                                       C.POINTER(C.c_double), fp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _sig_done = True
    return lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Net:
    """Device U-Net on an (ni, nj, nk) grid (multiples of 16) and `batch` samples."""

    def __init__(self, shape, batch=1, trainable=False, weights=None):
        self.lib = _lib()
        self.shape = tuple(int(s) for s in shape)
        self.batch = int(batch)
        h = C.c_void_p()
        st = self.lib.ls_net_create(C.byref(h), *self.shape, self.batch, 1 if trainable else 0)
        if st != 0:
            msg = self.lib.ls_net_last_error(h).decode() if h.value else ""
            if h.value:
                self.lib.ls_net_destroy(h)
            raise RuntimeError(f"ls_net_create failed ({st}): {msg}")
        self.h = h
        self.nw = int(self.lib.ls_net_num_weights())
        assert self.nw == S.n_weights()
        self.set_weights(S.init_weights(0) if weights is None else weights)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ls_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            raise RuntimeError(f"status {st}: {self.lib.ls_net_last_error(self.h).decode()}")

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._check(self.lib.ls_net_set_weights(self.h, _fp(w), w.size))

    def get_weights(self):
        w = np.empty(self.nw, np.float32)
        self._check(self.lib.ls_net_get_weights(self.h, _fp(w), w.size))
        return w

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == (self.batch,) + self.shape + (S.CIN,)
        out = np.empty((self.batch,) + self.shape, np.float32)
        self._check(self.lib.ls_net_forward(self.h, _fp(x), _fp(out)))
        return out

    def backward(self, gout):
        g = np.ascontiguousarray(gout, dtype=np.float32)
        gw = np.empty(self.nw, np.float32)
        self._check(self.lib.ls_net_backward(self.h, _fp(g), _fp(gw)))
        return gw

    def step(self, gout, lr):
        g = np.ascontiguousarray(gout, dtype=np.float32)
        self._check(self.lib.ls_net_step(self.h, _fp(g), float(lr)))

    def set_conv(self, tiled):
        """Convolution kernels: LDS-tiled (True, default) or the general
        implicit-GEMM kernel (False); both give the same bits."""
        self._check(self.lib.ls_net_set_conv(self.h, 1 if tiled else 0))

    def time_forward(self, reps=5):
        ms = C.c_double()
        self._check(self.lib.ls_net_time_forward(self.h, int(reps), C.byref(ms)))
        return ms.value

    # --- Galerkin energy training with the samples held on the device ---
    def store_alloc(self, samples, operators):
        self._check(self.lib.ls_net_store_alloc(self.h, int(samples), int(operators)))

    def store_operator(self, slot, ac):
        a = np.ascontiguousarray(ac, dtype=np.float32)
        assert a.shape == (4,) + self.shape
        self._check(self.lib.ls_net_store_operator(self.h, int(slot), _fp(a)))

    def store_sample(self, slot, op_slot, x, rc, eae):
        x = np.ascontiguousarray(x, dtype=np.float32)
        rc = np.ascontiguousarray(rc, dtype=np.float32)
        assert x.shape == self.shape + (S.CIN,) and rc.shape == self.shape
        self._check(self.lib.ls_net_store_sample(self.h, int(slot), int(op_slot), _fp(x), _fp(rc),
                                                 float(eae)))

    def energy_step(self, slots, lr, grad=False):
        """One batch of stored samples: losses (and, with grad and lr = 0, the
        weight gradient of their mean); lr > 0 takes an Adam step."""
        sl = (C.c_int * self.batch)(*[int(q) for q in slots])
        losses = (C.c_double * self.batch)()
        gw = np.empty(self.nw, np.float32) if grad else None
        self._check(self.lib.ls_net_energy_step(self.h, sl, float(lr), losses,
                                                _fp(gw) if grad else None))
        return np.array(losses[:]), gw

    def train_step(self, x, target, mask, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.ls_net_train_step(self.h, _fp(x), _fp(t), _fp(m), float(lr),
                                               C.byref(loss)))
        return loss.value
