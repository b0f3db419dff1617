"""ctypes binding of the CPU FP64 oracle (liblsoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
cpu_baseline leg of bench.py, as the values the GPU path is compared against
and the timed CPU baseline.  The product path (leachingsim_amd) never imports
this module.

``OracleSimulation`` exposes the same methods as
``leachingsim_amd.lsgpu.GpuSimulation`` so one test sequence drives both.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from leachingsim_amd.lsgpu import (F_FRAC, F_IMG, FIELD_DTYPES, PH_LEACHING, PH_PRESSURE,
                                   PH_TRANSPORT, STATE_FIELDS, Params, StepReport,
                                   default_params)

HERE = Path(__file__).resolve().parent
LIB = HERE / "liblsoracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liblsoracle.so"], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    sig = {
        "lso_create": (i, [C.POINTER(vp), i, i, i, C.POINTER(Params)]),
        "lso_destroy": (None, [vp]),
        "lso_last_error": (C.c_char_p, [vp]),
        "lso_set_params": (i, [vp, C.POINTER(Params)]),
        "lso_upload": (i, [vp, i, vp, C.c_size_t]),
        "lso_download": (i, [vp, i, vp, C.c_size_t]),
        "lso_init_state": (i, [vp, C.POINTER(ll)]),
        "lso_step": (i, [vp, d, d, C.POINTER(StepReport)]),
        "lso_phase": (i, [vp, i, d, d, C.POINTER(StepReport)]),
        "lso_snapshot": (i, [vp]),
        "lso_rollback": (i, [vp]),
        "lso_set_threads": (None, [i]),
        "lso_set_ksp": (i, [vp, i]),
        "lso_pressure_system": (ll, [vp, d, vp, vp, vp, vp]),
        "lso_transport_system": (ll, [vp, d, i, vp, vp, vp, vp]),
        "lso_saturation_system": (ll, [vp, d, vp, vp, vp, vp]),
        "lso_apply_pressure": (i, [vp, vp, vp]),
        "lso_pressure_classes": (i, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def set_threads(n: int) -> None:
    load().lso_set_threads(int(n))


class OracleSimulation:
    def __init__(self, shape, params: Params | None = None, label_bytes: int = 1):
        self.lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.N = int(np.prod(self.shape))
        self.label_bytes = label_bytes
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        st = self.lib.lso_create(C.byref(h), *self.shape, C.byref(self.params))
        if st != 0:
            raise RuntimeError(f"lso_create failed ({st})")
        self.h = h
        self.initial_sulphide_count = 0
        self.total_leached_count = 0

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.lso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow_positive=False):
        if st < 0 or (st > 0 and not allow_positive):
            raise RuntimeError(f"oracle status {st}: {self.lib.lso_last_error(self.h).decode()}")
        return st

    def set_ksp(self, which: int):
        """0: Jacobi-PCG pressure on the reduced system and BiCGStab transport
        (default); 1: the reference's GMRES(30) + left Jacobi for both,
        pressure on the full reference-form system, each solve issued twice."""
        self._check(self.lib.lso_set_ksp(self.h, which))

    def set_params(self, params: Params):
        self._check(self.lib.lso_set_params(self.h, C.byref(params)))
        self.params = params

    def upload(self, field, arr):
        if field == F_IMG:
            dt = np.uint8 if self.label_bytes == 1 else np.uint16
        else:
            dt = FIELD_DTYPES.get(field, np.float64)
        a = np.ascontiguousarray(arr, dtype=dt)
        assert a.shape == self.shape
        self._check(self.lib.lso_upload(self.h, field, a.ctypes.data, a.itemsize))

    def download(self, field):
        if field == F_IMG:
            dt = np.uint8 if self.label_bytes == 1 else np.uint16
        else:
            dt = FIELD_DTYPES.get(field, np.float64)
        a = np.empty(self.shape, dtype=dt)
        self._check(self.lib.lso_download(self.h, field, a.ctypes.data, a.itemsize))
        return a

    def state(self):
        return {f: self.download(f) for f in STATE_FIELDS}

    def read_labels(self, img):
        self.upload(F_IMG, img)

    def init_state(self):
        n = C.c_longlong()
        self._check(self.lib.lso_init_state(self.h, C.byref(n)))
        self.initial_sulphide_count = n.value
        return n.value

    def step(self, t, dt):
        rep = StepReport()
        st = self.lib.lso_step(self.h, t, dt, C.byref(rep))
        self._check(st, allow_positive=True)
        if st == 0:
            self.total_leached_count += rep.leached
        return rep

    def phase(self, ph, t=0.0, dt=1.0):
        rep = StepReport()
        st = self.lib.lso_phase(self.h, ph, t, dt, C.byref(rep))
        self._check(st, allow_positive=True)
        rep.status = st
        return rep

    def saveState(self):
        self._check(self.lib.lso_snapshot(self.h))

    def restoreState(self):
        self._check(self.lib.lso_rollback(self.h))

    # reference-form systems (for dense-solve pinning on tiny grids)
    def pressure_system(self, t):
        nnz = self.lib.lso_pressure_system(self.h, t, None, None, None, None)
        r = np.empty(nnz, np.int32); c = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64); b = np.empty(self.N, np.float64)
        self.lib.lso_pressure_system(self.h, t, r.ctypes.data, c.ctypes.data, v.ctypes.data,
                                     b.ctypes.data)
        return r, c, v, b

    def saturation_system(self, dt):
        nnz = self.lib.lso_saturation_system(self.h, dt, None, None, None, None)
        r = np.empty(nnz, np.int32); c = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64); b = np.empty(self.N, np.float64)
        self.lib.lso_saturation_system(self.h, dt, r.ctypes.data, c.ctypes.data, v.ctypes.data,
                                       b.ctypes.data)
        return r, c, v, b

    def transport_system(self, dt, field):
        nnz = self.lib.lso_transport_system(self.h, dt, field, None, None, None, None)
        r = np.empty(nnz, np.int32); c = np.empty(nnz, np.int32)
        v = np.empty(nnz, np.float64); b = np.empty(self.N, np.float64)
        self.lib.lso_transport_system(self.h, dt, field, r.ctypes.data, c.ctypes.data,
                                      v.ctypes.data, b.ctypes.data)
        return r, c, v, b

    def apply_pressure(self, x):
        a = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(a)
        self._check(self.lib.lso_apply_pressure(self.h, a.ctypes.data, y.ctypes.data))
        return y

    def pressure_classes(self):
        out = np.empty(self.shape, np.int8)
        self._check(self.lib.lso_pressure_classes(self.h, out.ctypes.data))
        return out


def dense(rows, cols, vals, n):
    """Dense matrix from COO, summing duplicates (PETSc ADD_VALUES)."""
    A = np.zeros((n, n))
    np.add.at(A, (rows, cols), vals)
    return A
