"""ctypes binding of the gfx950 hot-path library (include/lsgpu.h).

``GpuSimulation`` mirrors the per-step methods of the reference's
``class Simulation`` (src/Simulation.h:29-186): each method forwards to one
C-ABI call on a device-resident context.  There is no CPU fallback: if
``liblsgpu.so`` is missing or no HIP device is present the constructor
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("LSGPU_LIB", PKG / "liblsgpu.so"))

AIR, PORE, ROCK, SULPHIDE, PRECIPITATE = 1, 2, 3, 4, 5

SOLVER_PLAIN64, SOLVER_MIXED, SOLVER_LEARNED = 0, 1, 2
BC_REFERENCE, BC_ZFACES = 0, 1
WARM_PREVIOUS, WARM_PROJECTED, WARM_LEARNED = 0, 1, 2
TIME_RB_GATHER, TIME_RB_APPLY, TIME_SCHUR, TIME_TRED, TIME_TBLACK = 1, 2, 3, 4, 5
TIME_TRED32, TIME_TBLACK32 = 6, 7
WARM_MAX = 8

(F_IMG, F_FRAC, F_INVENTORY, F_SATURATION, F_PRESSURE, F_CONC, F_ACID, F_QX, F_QY, F_QZ,
 F_KREL, F_PC, F_PERM, F_MU) = range(14)

(PH_PROPERTIES, PH_PRESSURE, PH_FLUX, PH_SATURATION, PH_SURFACE, PH_TRANSPORT,
 PH_PRECIPITATION, PH_LEACHING) = range(8)

FIELD_DTYPES = {F_FRAC: np.float32}
STATE_FIELDS = (F_IMG, F_FRAC, F_INVENTORY, F_SATURATION, F_PRESSURE, F_CONC, F_ACID,
                F_QX, F_QY, F_QZ)


class Params(C.Structure):
    _fields_ = [
        ("D", C.c_double), ("dx", C.c_double), ("Dpore_fac", C.c_double),
        ("kext", C.c_double), ("kreac", C.c_double), ("theta", C.c_double),
        ("c_sat", C.c_double), ("evap_flux", C.c_double), ("top_pressure", C.c_double),
        ("max_cap_grad", C.c_double), ("vg_alpha", C.c_double), ("vg_n", C.c_double),
        ("s_res", C.c_double), ("porosity", C.c_double), ("solid_density", C.c_double),
        ("time_scalar", C.c_double), ("alpha_transfer", C.c_double),
        ("instant_precip", C.c_int), ("bc_mode", C.c_int), ("solver", C.c_int),
        ("warm_start", C.c_int), ("press_rtol", C.c_double), ("press_maxit", C.c_int),
        ("adr_rtol", C.c_double), ("adr_maxit", C.c_int), ("warm_history", C.c_int),
        ("mixed_inner_rtol", C.c_double), ("learned_race", C.c_double),
        ("implicit_sat", C.c_int), ("mixed_schur", C.c_int), ("adr_schur", C.c_int),
        ("adr_extrapolate", C.c_int), ("adr_mixed", C.c_int),
        ("adr_mixed_inner_rtol", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepReport(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("press_iters", C.c_int), ("press_outer", C.c_int),
        ("press_fallback", C.c_int), ("press_unet_applies", C.c_int),
        ("press_relres", C.c_double), ("press_bnorm", C.c_double),
        ("adr_iters", C.c_int), ("adr_relres", C.c_double),
        ("sat_substeps", C.c_int), ("sat_max_divq", C.c_double),
        ("leach_substeps", C.c_int), ("leach_max_rate", C.c_double),
        ("leached", C.c_int), ("leach_general_path", C.c_int),
        ("n_floored", C.c_longlong),
        ("p_min", C.c_double), ("p_max", C.c_double), ("conc_min", C.c_double),
        ("conc_max", C.c_double), ("acid_min", C.c_double), ("acid_max", C.c_double),
        ("mass_conc", C.c_double), ("mass_acid", C.c_double), ("mass_water", C.c_double),
        ("top_pressure_eff", C.c_double),
        ("t_props_ms", C.c_double), ("t_pressure_ms", C.c_double), ("t_flux_ms", C.c_double),
        ("t_sat_ms", C.c_double), ("t_surface_ms", C.c_double),
        ("t_transport_ms", C.c_double), ("t_precip_ms", C.c_double),
        ("t_leach_ms", C.c_double), ("t_snapshot_ms", C.c_double), ("t_total_ms", C.c_double),
        ("prof_stencil64_ms", C.c_double), ("prof_stencil32_ms", C.c_double),
        ("prof_stencil64_n", C.c_int), ("prof_stencil32_n", C.c_int),
        ("press_relres0", C.c_double), ("warm_basis", C.c_int), ("warm_net", C.c_int),
        ("sat_iters", C.c_int), ("sat_valid", C.c_int), ("sat_raw_min", C.c_double),
        ("sat_raw_max", C.c_double), ("prof_schur_ms", C.c_double), ("prof_schur_n", C.c_int),
        ("prof_schur_bytes", C.c_double), ("adr_iters_conc", C.c_int),
        ("adr_iters_acid", C.c_int), ("prof_tred_ms", C.c_double), ("prof_tred_n", C.c_int),
        ("prof_tred_bytes", C.c_double), ("adr_cycles", C.c_int), ("adr_fallback", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def default_params(**overrides) -> Params:
    """The reference's CLI defaults (src/main.cpp:39, 89-91), plus solver settings."""
    p = Params(D=1.0, dx=1.0, Dpore_fac=1.0, kext=1.0, kreac=1.0, theta=1.0, c_sat=0.95,
               evap_flux=0.0, top_pressure=10000.0, max_cap_grad=50000.0, vg_alpha=1.0,
               vg_n=2.0, s_res=0.05, porosity=0.1, solid_density=1.0, time_scalar=1.0,
               alpha_transfer=1e-5, instant_precip=0, bc_mode=BC_REFERENCE,
               solver=SOLVER_PLAIN64, warm_start=WARM_PREVIOUS, press_rtol=1e-8,
               press_maxit=10000, adr_rtol=1e-8, adr_maxit=10000, warm_history=1,
               mixed_inner_rtol=0.1, learned_race=1.0, implicit_sat=0,
               mixed_schur=1, adr_schur=1, adr_extrapolate=1, adr_mixed=1,
               adr_mixed_inner_rtol=1e-6)
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


_lib = None


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Load liblsgpu.so and declare its signatures.  Raises if missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} not built: run __graft_entry__.build() (the HIP path has no "
                           "CPU fallback)")
    lib = C.CDLL(str(p))
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    sig = {
        "ls_default_params": (None, [C.POINTER(Params)]),
        "ls_abi_info": (i, [C.POINTER(sz), C.POINTER(sz)]),
        "ls_device_bytes": (i, [vp, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
        "ls_create": (i, [C.POINTER(vp), i, i, i, i, C.POINTER(Params)]),
        "ls_destroy": (None, [vp]),
        "ls_last_error": (C.c_char_p, [vp]),
        "ls_get_params": (i, [vp, C.POINTER(Params)]),
        "ls_set_params": (i, [vp, C.POINTER(Params)]),
        "ls_upload": (i, [vp, i, vp, sz]),
        "ls_download": (i, [vp, i, vp, sz]),
        "ls_init_state": (i, [vp, C.POINTER(C.c_longlong)]),
        "ls_step": (i, [vp, d, d, C.POINTER(StepReport)]),
        "ls_phase": (i, [vp, i, d, d, C.POINTER(StepReport)]),
        "ls_snapshot": (i, [vp]),
        "ls_rollback": (i, [vp]),
        "ls_debug_apply_pressure": (i, [vp, vp, vp]),
        "ls_time_stencil": (i, [vp, i, i, C.POINTER(d), C.POINTER(d)]),
        "ls_time_copy": (i, [vp, sz, i, C.POINTER(d)]),
        "ls_set_profiling": (i, [vp, i]),
        "ls_set_network": (i, [vp, C.POINTER(C.c_float), sz]),
        "ls_network_grid": (i, [vp, C.POINTER(C.c_int)]),
        "ls_debug_learned_pack": (i, [vp, vp, vp, sz]),
        "ls_debug_learned_galerkin": (i, [vp, vp, sz]),
        "ls_debug_learned_sample": (i, [vp, vp, vp, sz, vp, C.POINTER(d)]),
        "ls_debug_pressure_system": (i, [vp, vp, vp, vp]),
        "ls_debug_pressure_faces": (i, [vp, vp]),
        "ls_debug_schur_apply": (i, [vp, i, vp, vp, C.POINTER(d)]),
        "ls_update_saturation_noalloc": (i, [i, i, i, vp, vp, vp, vp, vp, d, d, d, d, vp,
                                             C.POINTER(i), C.POINTER(d)]),
        "ls_noalloc_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    sp, sr = C.c_size_t(), C.c_size_t()
    lib.ls_abi_info(C.byref(sp), C.byref(sr))
    if (sp.value, sr.value) != (C.sizeof(Params), C.sizeof(StepReport)):
        raise RuntimeError(f"{p}: struct sizes {sp.value}/{sr.value} differ from this binding's "
                           f"{C.sizeof(Params)}/{C.sizeof(StepReport)} (rebuild the library)")
    if path is None:
        _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Names declared in include/lsgpu.h (for the export check)."""
    import re
    hdr = (PKG.parent / "include" / "lsgpu.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|const char\*)\s+(ls_\w+)\(", hdr, re.M)))


def update_saturation_noalloc(shape, img_ptr: int, qx_ptr: int, qy_ptr: int, qz_ptr: int,
                              s_ptr: int, dt: float, dx: float = 1.0, porosity: float = 0.1,
                              evap_flux: float = 0.0, stream: int | None = None,
                              read_plan: bool = True):
    """updateSaturation on caller-owned device fields (raw device addresses:
    uint8 labels, FP64 fluxes and saturation, k-fastest) without allocating
    device memory (ls_update_saturation_noalloc).  Returns (n_sub, max rate)
    when read_plan, else None (the call then stays asynchronous on `stream`)."""
    lib = load_library()
    ni, nj, nk = (int(x) for x in shape)
    ns, mx = C.c_int(0), C.c_double(0.0)
    st = lib.ls_update_saturation_noalloc(ni, nj, nk, img_ptr, qx_ptr, qy_ptr, qz_ptr, s_ptr,
                                          dt, dx, porosity, evap_flux, stream,
                                          C.byref(ns) if read_plan else None,
                                          C.byref(mx) if read_plan else None)
    if st != 0:
        raise LsError(st, lib.ls_noalloc_last_error().decode())
    return (ns.value, mx.value) if read_plan else None


class LsError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class GpuSimulation:
    """Device-resident simulation context (one HIP stream).

    Grid shape is (ni, nj, nk) = the reference's (i, j, k) = CLI (zext, yext,
    xext); arrays are numpy with that shape, k fastest.
    """

    def __init__(self, shape, params: Params | None = None, label_bytes: int = 1):
        self.lib = load_library()
        self.shape = tuple(int(s) for s in shape)
        self.label_bytes = label_bytes
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        st = self.lib.ls_create(C.byref(h), *self.shape, label_bytes, C.byref(self.params))
        if st != 0:
            msg = self.lib.ls_last_error(h).decode() if h.value else "ls_create failed"
            if h.value:
                self.lib.ls_destroy(h)
            raise LsError(st, msg)
        self.h = h
        self.initial_sulphide_count = 0
        self.total_leached_count = 0

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.ls_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st: int, allow_positive: bool = False) -> int:
        if st < 0 or (st > 0 and not allow_positive):
            raise LsError(st, self.lib.ls_last_error(self.h).decode())
        return st

    def set_params(self, params: Params):
        self._check(self.lib.ls_set_params(self.h, C.byref(params)))
        self.params = params

    # -- fields -------------------------------------------------------------
    def upload(self, field: int, arr: np.ndarray):
        if field == F_IMG:
            dt = np.uint8 if self.label_bytes == 1 else np.uint16
        else:
            dt = FIELD_DTYPES.get(field, np.float64)
        a = np.ascontiguousarray(arr, dtype=dt)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        self._check(self.lib.ls_upload(self.h, field, a.ctypes.data, a.itemsize))

    def download(self, field: int) -> np.ndarray:
        if field == F_IMG:
            dt = np.uint8 if self.label_bytes == 1 else np.uint16
        else:
            dt = FIELD_DTYPES.get(field, np.float64)
        a = np.empty(self.shape, dtype=dt)
        self._check(self.lib.ls_download(self.h, field, a.ctypes.data, a.itemsize))
        return a

    def state(self) -> dict:
        return {f: self.download(f) for f in STATE_FIELDS}

    # -- reference initialisation (src/main.cpp:189-203) ----------------------
    def read_labels(self, img: np.ndarray):
        self.upload(F_IMG, img)

    def init_state(self) -> int:
        n = C.c_longlong()
        self._check(self.lib.ls_init_state(self.h, C.byref(n)))
        self.initial_sulphide_count = n.value
        self.total_leached_count = 0
        return n.value

    # -- the coupled step and its phases --------------------------------------
    def step(self, t: float, dt: float) -> StepReport:
        rep = StepReport()
        st = self.lib.ls_step(self.h, t, dt, C.byref(rep))
        self._check(st, allow_positive=True)
        if st == 0:
            self.total_leached_count += rep.leached
        return rep

    def phase(self, ph: int, t: float = 0.0, dt: float = 1.0) -> StepReport:
        rep = StepReport()
        st = self.lib.ls_phase(self.h, ph, t, dt, C.byref(rep))
        self._check(st, allow_positive=True)
        return rep

    # reference method names (src/Simulation.h:42-70)
    def updateProperties(self):
        return self.phase(PH_PROPERTIES)

    def solvePressure(self, t: float) -> bool:
        return self.phase(PH_PRESSURE, t).status == 0

    def calculateFlux(self):
        return self.phase(PH_FLUX)

    def updateSaturation(self, dt: float):
        return self.phase(PH_SATURATION, 0.0, dt)

    def handleSurfaceEffects(self):
        return self.phase(PH_SURFACE)

    def solveTransport(self, dt: float) -> bool:
        return self.phase(PH_TRANSPORT, 0.0, dt).status == 0

    def handlePrecipitation(self):
        return self.phase(PH_PRECIPITATION)

    def updateFrac(self, dt: float) -> int:
        return self.phase(PH_LEACHING, 0.0, dt).leached

    def saveState(self):
        self._check(self.lib.ls_snapshot(self.h))

    def restoreState(self):
        self._check(self.lib.ls_rollback(self.h))

    # -- test / measurement hooks ---------------------------------------------
    def apply_pressure(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(a)
        self._check(self.lib.ls_debug_apply_pressure(self.h, a.ctypes.data, y.ctypes.data))
        return y

    def time_stencil(self, precision: int = 64, reps: int = 20):
        """ms per launch and algorithmic bytes per launch of one kernel:
        precision 64 / 32 (the stencil), TIME_RB_GATHER, TIME_RB_APPLY,
        TIME_SCHUR (pressure), TIME_TRED, TIME_TBLACK (the transport's
        Schur passes on the rows of the last transport solve), TIME_TRED32,
        TIME_TBLACK32 (the same passes of the mixed-precision form)."""
        ms, by = C.c_double(), C.c_double()
        self._check(self.lib.ls_time_stencil(self.h, precision, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_network(self, weights):
        w = np.ascontiguousarray(weights, dtype=np.float32)
        self._check(self.lib.ls_set_network(self.h, w.ctypes.data_as(C.POINTER(C.c_float)), w.size))
        self._net_weights = w

    def network_grid(self):
        d = (C.c_int * 3)()
        self._check(self.lib.ls_network_grid(self.h, d))
        return tuple(d)

    def learned_pack(self, r):
        a = np.ascontiguousarray(r, dtype=np.float32)
        dims = self.network_grid()
        out = np.empty(dims + (16,), np.float32)
        self._check(self.lib.ls_debug_learned_pack(self.h, a.ctypes.data, out.ctypes.data, out.size))
        return out

    def learned_galerkin(self):
        """Galerkin operator P^T A^ P of the assembled scaled pressure operator
        on the network grid: [4, ci, cj, ck] float32 (diagonal, couplings to
        the +i, +j, +k neighbours)."""
        dims = self.network_grid()
        out = np.empty((4,) + dims, np.float32)
        self._check(self.lib.ls_debug_learned_galerkin(self.h, out.ctypes.data, out.size))
        return out

    def learned_sample(self, p_star):
        """(network input [ci, cj, ck, 16], rc = P^T r [ci, cj, ck], e.A^ e) for
        e = sqrt(d) (p_star - p) on the solved rows, p the current pressure
        field, r = A^ e."""
        ps = np.ascontiguousarray(p_star, dtype=np.float64)
        assert ps.shape == self.shape
        dims = self.network_grid()
        x = np.empty(dims + (16,), np.float32)
        rc = np.empty(dims, np.float32)
        eae = C.c_double()
        self._check(self.lib.ls_debug_learned_sample(self.h, ps.ctypes.data, x.ctypes.data,
                                                     x.size, rc.ctypes.data, C.byref(eae)))
        return x, rc, eae.value

    def pressure_system(self):
        sd = np.empty(self.shape)
        bh = np.empty(self.shape)
        cl = np.empty(self.shape, np.int8)
        self._check(self.lib.ls_debug_pressure_system(self.h, sd.ctypes.data, bh.ctypes.data,
                                                      cl.ctypes.data))
        return sd, bh, cl

    def pressure_faces(self):
        f = np.empty((3,) + self.shape)
        self._check(self.lib.ls_debug_pressure_faces(self.h, f.ctypes.data))
        return f

    def schur_apply(self, p_black, form: int = 1):
        """q = S p on a vector compacted to the black voxels (N/2 float32,
        black voxel v at v >> 1) with the dot p.q; form 1 fused, 2 two-pass;
        form 4: two-pass on a full-length vector (N float32, any nk; red
        entries ignored on input, zero on output)."""
        a = np.ascontiguousarray(p_black, dtype=np.float32)
        n = int(np.prod(self.shape))
        assert a.size == (n if form == 4 else n // 2) and (form == 4 or n % 2 == 0)
        q = np.empty_like(a)
        dot = C.c_double()
        self._check(self.lib.ls_debug_schur_apply(self.h, form, a.ctypes.data, q.ctypes.data,
                                                  C.byref(dot)))
        return q, dot.value

    def set_profiling(self, on: bool):
        self._check(self.lib.ls_set_profiling(self.h, 1 if on else 0))

    def device_bytes(self) -> tuple[int, int]:
        """(grid-proportional, other) device bytes the context holds."""
        g, o = C.c_ulonglong(), C.c_ulonglong()
        self._check(self.lib.ls_device_bytes(self.h, C.byref(g), C.byref(o)))
        return g.value, o.value

    def time_copy(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double()
        self._check(self.lib.ls_time_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value
