"""The configured workloads' geometry classes on the GPU.

* The 512^3 workload's class (uint16 labels: pore + Rock / Sulphide /
  Precipitate at 0.30 / 0.25 / 0.15 from their own seeded fields, a 4-voxel
  planar fracture, 2-plane Air slabs on the j faces, z-face inlet/outlet,
  dt 0.5) at 48^3 and 64^3, 10 coupled steps against the CPU oracle with the
  mixed solver + projected warm start and with the learned solver + learned
  warm start.  At rtol 1e-12 both sides' solution errors are ~cond * 1e-12;
  this class has the largest conductance contrast of the workloads (Sulphide
  and Precipitate next to Pore), and p must agree to 1e-8 (the required
  bar against the FP64 oracle), conc to 1e-10; masks are bit-exact every
  step.
* The 338x338x283 workload (uint8, sigma 3, porosity 0.30, z-face
  inlet/outlet) at full size, where the oracle is too slow: 3 steps with
  plain FP64 PCG and with the mixed solver, each converged to 1e-8 in FP64,
  equal masks, finite mass sums, and pressures that agree within what two
  1e-8 solves allow.
"""
import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L
from tests.helpers import both, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,solver", [(48, L.SOLVER_MIXED), (64, L.SOLVER_MIXED),
                                      (48, L.SOLVER_LEARNED)])
def test_fractured_class_matches_oracle(n, solver):
    img = G.named_geometry("fractured512", size=n)
    assert img.dtype == np.uint16
    warm = L.WARM_LEARNED if solver == L.SOLVER_LEARNED else L.WARM_PROJECTED
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=solver, warm_start=warm,
                         press_rtol=1e-12, adr_rtol=1e-12)
    g, o = both(img.shape, P, img, label_bytes=2)
    if solver == L.SOLVER_LEARNED:
        from leachingsim_amd.learn import load_weights
        g.set_network(load_weights())
    print(f"fractured {n}^3 solver {solver}", flush=True)
    t, dt = 0.0, 0.5
    worst = [0.0, 0.0]
    changed = 0
    for step in range(10):
        rg = g.step(t, dt)
        ro = o.step(t, dt)
        assert rg.status == 0 and ro.status == 0, (step, rg.status, ro.status)
        assert rg.press_relres <= 1e-12 and rg.adr_relres <= 1e-12
        lab = g.download(L.F_IMG)
        assert lab.dtype == np.uint16
        assert np.array_equal(lab, o.download(L.F_IMG)), step
        assert np.array_equal(g.download(L.F_FRAC), o.download(L.F_FRAC)), step
        changed = int((lab != img).sum())
        ep = relerr(g.download(L.F_PRESSURE), o.download(L.F_PRESSURE))
        ec = relerr(g.download(L.F_CONC), o.download(L.F_CONC))
        worst = [max(worst[0], ep), max(worst[1], ec)]
        assert ep <= 1e-8 and ec <= 1e-10, (step, ep, ec)
        t += dt
    for f in (L.F_SATURATION, L.F_INVENTORY, L.F_QX, L.F_QY, L.F_QZ):
        assert relerr(g.download(f), o.download(f)) <= 1e-8, f
    print(f"  worst |dp|/|p| {worst[0]:.2e}, |dc|/|c| {worst[1]:.2e}, "
          f"labels changed {changed}")
    # surface precipitation at the Air slabs is active in this class
    assert changed > 0
    g.close()


def test_smooth338_full_size_properties():
    img = G.named_geometry("smooth338")
    assert img.shape == (283, 338, 338)
    runs = {}
    for solver in (L.SOLVER_PLAIN64, L.SOLVER_MIXED):
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=solver, warm_start=L.WARM_PROJECTED)
        g = L.GpuSimulation(img.shape, P)
        g.read_labels(img)
        g.init_state()
        print(f"smooth338 solver {solver}", flush=True)
        t, reps = 0.0, []
        for step in range(3):
            r = g.step(t, 1.0)
            assert r.status == 0, step
            assert r.press_relres <= 1e-8 and r.adr_relres <= 1e-8
            for x in (r.mass_conc, r.mass_acid, r.mass_water, r.p_min, r.p_max):
                assert np.isfinite(x)
            reps.append((r.press_iters, r.t_total_ms))
            t += 1.0
        print(f"  iterations, ms per step: {reps}")
        runs[solver] = (g.download(L.F_IMG), g.download(L.F_PRESSURE), reps)
        g.close()
    (img64, p64, r64), (img32, p32, r32) = runs[L.SOLVER_PLAIN64], runs[L.SOLVER_MIXED]
    assert np.array_equal(img64, img32)
    assert np.all(np.isfinite(p64)) and np.all(np.isfinite(p32))
    assert p64.max() > 0
    err = relerr(p32, p64)
    print(f"  mixed vs plain FP64: |dp|/|p| {err:.2e}")
    # two solves stopped at 1e-8 differ by up to cond * 1e-8
    assert err <= 1e-5


def test_device_bytes_per_voxel_of_the_1024x1024x768_class():
    """The 1024x1024x768 workload's class at 256x256x192 with the mixed
    solver and the projected warm start: after three steps the context holds
    at most 330 B per voxel in grid-sized arrays (state, snapshot,
    properties, the phase pool, history), so the full grid (805M voxels)
    fits one 288 GiB MI355X with the runtime's own use beside it."""
    img = G.named_geometry("fractured1024x768", size=256)
    assert img.shape == (192, 256, 256) and img.dtype == np.uint8
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED,
                         warm_start=L.WARM_PROJECTED)
    g = L.GpuSimulation(img.shape, P)
    g.read_labels(img)
    g.init_state()
    print("fractured1024x768 class at 256x256x192", flush=True)
    t = 0.0
    for step in range(3):
        r = g.step(t, 0.5)
        assert r.status == 0 and r.press_relres <= 1e-8 and r.adr_relres <= 1e-8, step
        t += 0.5
    grid, other = g.device_bytes()
    per_voxel = grid / img.size
    print(f"  grid bytes {grid / 2**30:.2f} GiB = {per_voxel:.1f} B/voxel, other "
          f"{other / 2**20:.1f} MiB")
    assert 280.0 <= per_voxel <= 330.0
    g.close()


def test_more_row_blocks_than_one_grid_dimension():
    """A grid of 1040 x 1040 rows (nk 8): 67,600 blocks of 16 rows, more
    than one grid dimension holds, so the row kernels (flux, saturation and
    transport rows with a voxel pair per lane, and the transport's Schur
    passes in FP64 and FP32) fold their row blocks over grid y and z.  Each
    phase against the oracle on the same input: fluxes to 1e-13, saturation
    bit for bit, and the mixed-precision transport to 1e-10 at adr_rtol
    1e-12."""
    from tests.helpers import small_mixed_geometry
    shape = (1040, 1040, 8)
    img = small_mixed_geometry(shape, 31)
    P = L.default_params(bc_mode=L.BC_ZFACES, press_rtol=1e-12, adr_rtol=1e-12,
                         evap_flux=1e-7)
    g, o = both(shape, P, img)
    rng = np.random.default_rng(7)
    print(f"rows {shape[0] * shape[1]}, voxels {img.size}", flush=True)
    p = rng.uniform(0.0, 1e4, size=shape)
    s = rng.uniform(0.0, 1.0, size=shape)
    for x in (g, o):
        x.upload(L.F_PRESSURE, p)
        x.upload(L.F_SATURATION, s)
        x.phase(L.PH_PROPERTIES)
        x.phase(L.PH_FLUX)
    for f in (L.F_QX, L.F_QY, L.F_QZ):
        # against the field's scale: among 8.6M voxels some fluxes come from
        # near-cancelling gradients, where a rounding step is a large part
        # of the value
        a, b = g.download(f), o.download(f)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)), f
    # the same fluxes on both sides from here: saturation is then bit-exact
    for f in (L.F_QX, L.F_QY, L.F_QZ):
        g.upload(f, o.download(f))
    rg = g.phase(L.PH_SATURATION, 0.0, 5.0)
    ro = o.phase(L.PH_SATURATION, 0.0, 5.0)
    assert rg.sat_substeps == ro.sat_substeps
    assert np.array_equal(g.download(L.F_SATURATION), o.download(L.F_SATURATION))
    # transport on the fluxes both sides hold, conc from a random field
    conc = rng.uniform(0.0, 1.0, size=shape)
    for x in (g, o):
        x.upload(L.F_CONC, conc)
    rg = g.phase(L.PH_TRANSPORT, 0.0, 0.5)
    ro = o.phase(L.PH_TRANSPORT, 0.0, 0.5)
    assert rg.status == 0 and ro.status == 0
    assert rg.adr_cycles >= 1 and rg.adr_fallback == 0
    assert relerr(g.download(L.F_CONC), o.download(L.F_CONC)) <= 1e-10
    for f in (L.F_QX, L.F_QY, L.F_QZ):   # zeroed at Rock / Precipitate by the row builder
        assert np.array_equal(g.download(f), o.download(f)), f
