"""The implicit saturation solve of the oracle (setupSaturationEqns +
solveSaturation, src/Simulation.cpp:1858-2053) against a dense solve of its
assembly, closed forms that pin the assembly itself, and the physics check
that makes the reference roll the step back."""
import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L
from oracle.oracle import OracleSimulation, dense


def make(shape, img, **kw):
    P = L.default_params(implicit_sat=1, adr_rtol=1e-13, **kw)
    o = OracleSimulation(shape, P)
    o.read_labels(img)
    o.init_state()
    return o


def test_matches_dense_solve_of_its_assembly():
    shape = (5, 6, 4)
    rng = np.random.default_rng(0)
    img = rng.choice([G.PORE, G.PORE, G.PORE, G.ROCK, G.AIR, G.SULPHIDE], size=shape).astype(np.uint8)
    o = make(shape, img, evap_flux=1e-4, porosity=0.3)
    s0 = np.where(img == G.PORE, 0.2 + 0.6 * rng.random(shape), rng.random(shape))
    o.upload(L.F_SATURATION, s0)
    for f in (L.F_QX, L.F_QY, L.F_QZ):
        o.upload(f, 1e-3 * rng.standard_normal(shape))
    r, c, v, b = o.saturation_system(2.0)
    x = np.linalg.solve(dense(r, c, v, s0.size), b)
    rep = o.phase(L.PH_SATURATION, 0.0, 2.0)
    assert rep.status == 0
    raw_ok = x.max() <= 1.05 and x.min() >= -0.05
    assert rep.sat_valid == int(raw_ok)
    expect = np.clip(x, 0, 1) if raw_ok else x
    np.testing.assert_allclose(o.download(L.F_SATURATION).ravel(), expect, rtol=1e-9, atol=1e-12)
    # inactive voxels are identity rows: unchanged
    inactive = img.ravel() != G.PORE
    assert np.array_equal(o.download(L.F_SATURATION).ravel()[inactive], s0.ravel()[inactive])


def test_zero_flux_evaporation_closed_form():
    """No flux: phi/dt s_new = phi/dt s - (#Air faces) evap / dx."""
    shape = (4, 3, 5)
    img = np.full(shape, G.PORE, np.uint8)
    img[0] = G.AIR
    phi, dt, ev = 0.25, 10.0, 2e-3
    o = make(shape, img, evap_flux=ev, porosity=phi)
    s0 = np.full(shape, 0.8)
    o.upload(L.F_SATURATION, s0)
    for f in (L.F_QX, L.F_QY, L.F_QZ):
        o.upload(f, np.zeros(shape))
    o.phase(L.PH_SATURATION, 0.0, dt)
    s = o.download(L.F_SATURATION)
    air_faces = np.zeros(shape)
    air_faces[1] = 1                                   # the plane next to the Air plane
    expect = np.where(img == G.PORE, s0 - dt * air_faces * ev / phi, s0)
    np.testing.assert_allclose(s, expect, rtol=1e-12)


def test_one_dimensional_upwind_recursion():
    """A column of Pore with uniform flux q > 0 along i (face flux = the mean
    of the two cells' flux = q), s0 uniform: voxel 0 only drains, each later
    voxel takes what the one upstream sends, the last one has no outflow."""
    ni = 9
    shape = (ni, 2, 2)
    img = np.full(shape, G.PORE, np.uint8)
    phi, dt, dx, q, s0 = 0.3, 5.0, 1.0, 2e-3, 0.5
    o = make(shape, img, porosity=phi)
    o.upload(L.F_SATURATION, np.full(shape, s0))
    o.upload(L.F_QX, np.full(shape, q))
    o.upload(L.F_QY, np.zeros(shape))
    o.upload(L.F_QZ, np.zeros(shape))
    rep = o.phase(L.PH_SATURATION, 0.0, dt)
    a = phi / dt
    w = q / s0 / dx
    expect = np.zeros(ni)
    expect[0] = a * s0 / (a + w)
    for i in range(1, ni - 1):
        expect[i] = (a * s0 + w * expect[i - 1]) / (a + w)
    expect[-1] = (a * s0 + w * expect[-2]) / a
    assert rep.sat_valid == 1
    for j in range(2):
        for k in range(2):
            np.testing.assert_allclose(o.download(L.F_SATURATION)[:, j, k], np.clip(expect, 0, 1),
                                       rtol=1e-11)


def test_unphysical_solution_is_flagged_and_not_clamped():
    ni = 6
    shape = (ni, 2, 2)
    img = np.full(shape, G.PORE, np.uint8)
    o = make(shape, img, porosity=0.1)
    o.upload(L.F_SATURATION, np.full(shape, 0.5))
    o.upload(L.F_QX, np.full(shape, 1.0))            # a huge flux piles up at the last voxel
    o.upload(L.F_QY, np.zeros(shape))
    o.upload(L.F_QZ, np.zeros(shape))
    rep = o.phase(L.PH_SATURATION, 0.0, 100.0)
    assert rep.sat_valid == 0 and rep.sat_raw_max > 1.05
    assert o.download(L.F_SATURATION).max() == pytest.approx(rep.sat_raw_max)


def test_picard_step_runs_and_rolls_back_when_unphysical():
    img = G.smoothed_threshold((10, 10, 12), 1.5, 0.45, 4)
    img[0] = G.AIR
    o = make(img.shape, img, bc_mode=L.BC_ZFACES)
    rep = o.step(0.0, 1.0)
    assert rep.status == 0 and rep.sat_valid == 1 and rep.sat_iters >= 0
