"""The allocation-free entry point ls_update_saturation_noalloc against the
saturation phase of a context (ls_phase, LS_PH_SATURATION), through the C
ABI: two calls in a row on caller-owned device buffers, the second on a
smaller grid (its scratch keeps the first call's contents), bit for bit,
with the same sub-step count and maximum rate; then the asynchronous form
(no plan read-back) on the first grid again.  The device buffers are the
caller's, allocated here through the HIP runtime the library links; the
library allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L

pytestmark = pytest.mark.gpu

PARAMS = dict(dx=1.3, porosity=0.25, evap_flux=2e-4)


def case(shape, seed):
    rng = np.random.default_rng(seed)
    img = rng.choice([G.AIR, G.PORE, G.PORE, G.PORE, G.ROCK, G.SULPHIDE, G.PRECIPITATE],
                     size=shape).astype(np.uint8)
    q = [1e-2 * rng.standard_normal(shape) for _ in range(3)]
    s = rng.uniform(0.0, 1.0, size=shape)
    return img, q, s


def with_context(img, q, s, dt):
    P = L.default_params(**PARAMS)
    g = L.GpuSimulation(img.shape, P)
    g.read_labels(img)
    g.init_state()
    for f, a in zip((L.F_QX, L.F_QY, L.F_QZ), q):
        g.upload(f, a)
    g.upload(L.F_SATURATION, s)
    r = g.phase(L.PH_SATURATION, 0.0, dt)
    assert r.status == 0
    out = (g.download(L.F_SATURATION), r.sat_substeps, r.sat_max_divq)
    g.close()
    return out


class Hip:
    """The caller's own device memory, through the HIP runtime (libamdhip64)."""
    H2D, D2H = 1, 2

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), a.nbytes) == 0
        self.held.append(p.value)
        assert self.lib.hipMemcpy(p.value, a.ctypes.data, a.nbytes, self.H2D) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def without_allocation(img, q, s, dt, read_plan=True):
    hip = Hip()
    try:
        pi = hip.put(img)
        pq = [hip.put(a) for a in q]
        ps = hip.put(s)
        plan = L.update_saturation_noalloc(img.shape, pi, pq[0], pq[1], pq[2], ps, dt,
                                           stream=None, read_plan=read_plan, **PARAMS)
        out = hip.get(ps, s)
    finally:
        hip.free()
    return out, plan


def test_noalloc_matches_context_phase_twice():
    dt = 3.0
    for n, (shape, seed) in enumerate((((40, 36, 48), 1), ((17, 22, 30), 2))):
        print(f"no-alloc call {n}: shape {shape}", flush=True)
        img, q, s = case(shape, seed)
        s_ref, nsub_ref, max_ref = with_context(img, q, s, dt)
        s_out, (nsub, mx) = without_allocation(img, q, s, dt)
        assert nsub == nsub_ref and nsub > 1
        assert mx == max_ref
        assert np.array_equal(s_out, s_ref)
        assert not np.array_equal(s_out, s)       # the update did something
    # asynchronous form (nothing read back), on the larger grid again
    img, q, s = case((40, 36, 48), 3)
    s_ref, _, _ = with_context(img, q, s, dt)
    s_out, plan = without_allocation(img, q, s, dt, read_plan=False)
    assert plan is None
    assert np.array_equal(s_out, s_ref)


def test_noalloc_rejects_sizes_beyond_its_limits():
    with pytest.raises(L.LsError, match="ni <= 65535"):
        L.update_saturation_noalloc((70000, 2, 2), 1, 1, 1, 1, 1, 1.0)
    with pytest.raises(L.LsError, match="2\\^32"):
        L.update_saturation_noalloc((65535, 65535, 2), 1, 1, 1, 1, 1, 1.0)
    with pytest.raises(L.LsError, match="null field"):
        L.update_saturation_noalloc((4, 4, 4), 0, 1, 1, 1, 1, 1.0)
