"""Pins the CPU oracle against closed forms, dense direct solves of the
reference-form systems, invariants and hand-built rule tables.

The reference ships no tests, fixtures or golden vectors (its *.raw files
are git-ignored and no logs are committed), so each oracle function is
checked against what the mathematics fixes, chosen so that a dropped term,
a wrong sign or index, or a transposed operand fails one of the checks.
"""
import math

import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L
from oracle.oracle import OracleSimulation, dense


def make(shape, img, **kw):
    o = OracleSimulation(shape, L.default_params(**kw))
    o.read_labels(img)
    o.init_state()
    return o


# ---------------------------------------------------------------------------
# Properties: van Genuchten with n = 2 (m = 1/2) has the closed forms
#   krel = sqrt(Se) (1 - sqrt(1 - Se^2))^2,   Pc = sqrt(Se^-2 - 1) / alpha
# ---------------------------------------------------------------------------
def vg2_krel(s):
    se = min(1.0, max(0.0, (s - 0.05) / 0.95))
    return 0.0 if se <= 0 else math.sqrt(se) * (1.0 - math.sqrt(1.0 - se * se)) ** 2


def vg2_pc(s):
    se = min(1.0, max(1e-6, (s - 0.05) / 0.95))
    return 0.0 if se >= 1.0 else math.sqrt(1.0 / (se * se) - 1.0)


@pytest.mark.parametrize("s", [0.0, 0.03, 0.05, 0.1, 0.37, 0.9, 0.999999, 1.0, 1.3])
def test_vg_closed_forms(s):
    shape = (2, 2, 3)
    img = np.array([G.PORE, G.SULPHIDE, G.AIR, G.ROCK, G.PRECIPITATE, G.PORE] * 2,
                   np.uint8).reshape(shape)
    o = make(shape, img)
    o.upload(L.F_SATURATION, np.full(shape, s))
    kr = o.download(L.F_KREL).ravel()
    pc = o.download(L.F_PC).ravel()
    for v, lab in enumerate(img.ravel()):
        if lab in (G.PORE, G.SULPHIDE):
            assert kr[v] == pytest.approx(vg2_krel(s), rel=1e-12, abs=1e-300)
            assert pc[v] == pytest.approx(vg2_pc(s), rel=1e-12)
        else:
            assert kr[v] == 0.0 and pc[v] == 0.0


def test_known_values_at_initial_saturation():
    # s = 0.1 (initSaturation) -> krel ~ 4.4070801e-7, Pc ~ 18.973666
    assert vg2_krel(0.1) == pytest.approx(4.4070801e-7, rel=1e-7)
    assert vg2_pc(0.1) == pytest.approx(18.9736660, rel=1e-8)


def test_permeability_and_viscosity_table():
    shape = (2, 2, 4)
    img = np.array([G.AIR, G.PORE, G.PORE, G.PORE, G.SULPHIDE, G.SULPHIDE, G.ROCK,
                    G.PRECIPITATE] * 2, np.uint8).reshape(shape)
    inv = np.array([0, 0.0, 1.0, 0.5, 0, 0, 0, 0] * 2, float).reshape(shape)
    frac = np.array([0, 0, 0, 0, 1.0, 0.0, 0, 0] * 2, np.float32).reshape(shape)
    o = make(shape, img)
    o.upload(L.F_SATURATION, np.full(shape, 1.0))   # krel = 1 on Pore/Sulphide
    o.upload(L.F_INVENTORY, inv)
    o.upload(L.F_FRAC, frac)
    K = o.download(L.F_PERM).ravel()[:8]
    mu = o.download(L.F_MU).ravel()[:8]
    expect = [1e-12 * 1e-15, 1e-12, 1e-18, 1e-15, 1e-18, 1e-12, 1e-18 * 1e-15, 1e-18 * 1e-15]
    for a, b in zip(K, expect):
        assert a == pytest.approx(b, rel=1e-13)
    assert list(mu) == [1.8e-5] + [1e-3] * 7


# ---------------------------------------------------------------------------
# Pressure: closed forms and dense solves of the reference-form system
# ---------------------------------------------------------------------------
def hmean(a, b):
    return 2 * a * b / (a + b)


def test_channel_series_resistance_closed_form():
    """An all-Pore (ni, 2, 2) bar, z-face planes: by symmetry the four columns
    carry equal pressure, so the solution is that of resistors in series
    (uniform pore links, K_air = 1e-9 end links).  Rock neighbours would leak
    through the 1e-25 permeability floor, so there are none."""
    ni = 17
    shape = (ni, 2, 2)
    img = np.full(shape, G.PORE, np.uint8)
    o = make(shape, img, bc_mode=L.BC_ZFACES, press_rtol=1e-14)
    o.phase(L.PH_PROPERTIES)
    t = 1000.0
    o.phase(L.PH_PRESSURE, t)
    p = o.download(L.F_PRESSURE)[:, 1, 1]
    assert np.ptp(o.download(L.F_PRESSURE), axis=(1, 2)).max() <= 1e-9 * p.max()
    ptop = 1e4 * t / 3600 + 1e-6
    K = 1e-12 * vg2_krel(0.1)
    c_mid = hmean(K, K) / 1e-3
    c_end = hmean(K, 1e-9) / 1e-3
    R = 2 / c_end + (ni - 1) / c_mid
    Q = ptop / R
    expect = ptop - Q / c_end - np.arange(ni) * Q / c_mid
    np.testing.assert_allclose(p, expect, rtol=1e-10)


def test_two_layers_harmonic_flux():
    """Half of the bar is partly clogged (inventory 0.1, K 10^-0.6 lower; at
    0.5 its rows would fall under the 1e-16 diagonal floor): the flux is the
    same through every link and equals ptop / (sum of resistances)."""
    ni = 12
    shape = (ni, 2, 2)
    img = np.full(shape, G.PORE, np.uint8)
    inv = np.zeros(shape)
    inv[ni // 2:] = 0.1
    o = make(shape, img, bc_mode=L.BC_ZFACES, press_rtol=1e-14)
    o.upload(L.F_INVENTORY, inv)
    o.phase(L.PH_PROPERTIES)
    o.phase(L.PH_PRESSURE, 3600.0)
    p = o.download(L.F_PRESSURE)[:, 1, 1]
    kr = vg2_krel(0.1)
    K = np.where(np.arange(ni) < ni // 2, 1e-12, 10.0 ** (-12.6)) * kr
    links = [hmean(K[0], 1e-9)] + [hmean(K[i], K[i + 1]) for i in range(ni - 1)] + \
        [hmean(K[-1], 1e-9)]
    c = np.array(links) / 1e-3
    Q = 1e4 / np.sum(1 / c)
    pots = np.concatenate([[1e4], p, [0.0]])
    np.testing.assert_allclose(c * -np.diff(pots), Q, rtol=1e-10)


def test_isolated_pore_is_floored_and_zero():
    shape = (5, 5, 5)
    img = np.full(shape, G.ROCK, np.uint8)
    img[2, 2, 2] = G.PORE
    img[0] = G.AIR
    o = make(shape, img)
    o.phase(L.PH_PROPERTIES)
    rep = o.phase(L.PH_PRESSURE, 100.0)
    assert rep.n_floored == 1
    assert o.download(L.F_PRESSURE)[2, 2, 2] == 0.0


@pytest.mark.parametrize("seed,bc", [(1, L.BC_REFERENCE), (2, L.BC_ZFACES), (3, L.BC_ZFACES)])
def test_pressure_matches_dense_solve_of_reference_system(seed, bc):
    shape = (6, 7, 8)
    img = G.smoothed_threshold(shape, 1.0, 0.55, seed)
    img[0] = G.AIR
    rng = np.random.default_rng(seed)
    img[(rng.random(shape) < 0.05) & (img == G.ROCK)] = G.SULPHIDE
    o = make(shape, img, bc_mode=bc, press_rtol=1e-14)
    rng_s = rng.uniform(0.08, 0.6, size=shape)
    o.upload(L.F_SATURATION, rng_s)
    o.phase(L.PH_PROPERTIES)
    t = 1234.5
    o.phase(L.PH_PRESSURE, t)
    p = o.download(L.F_PRESSURE).ravel()
    cls = o.pressure_classes().ravel()
    r, c, v, b = o.pressure_system(t)
    A = dense(r, c, v, p.size)
    x = np.linalg.solve(A, b)
    keep = cls != 2
    assert np.any(cls == 1)
    np.testing.assert_allclose(p[keep], x[keep], rtol=1e-9, atol=1e-9 * np.abs(x).max())
    # floored rows satisfy their own reference row exactly (post-solve)
    res = A @ p - b
    if np.any(cls == 2):
        fl = np.flatnonzero(cls == 2)
        assert np.all(np.abs(res[fl]) <= 1e-6 * np.abs(np.diag(A)[fl] * p[fl]).max() + 1e-12)


@pytest.mark.parametrize("rtol", [1e-6, 1e-13])
def test_reference_gmres_solver_meets_its_stop_test(rtol):
    """The reference's KSP setting (GMRES(30) + left Jacobi on the full
    reference-form system, src/Simulation.cpp:146-157): after the two solves
    the preconditioned residual recomputed from the assembled system meets
    ||D^-1 (b - A p)|| <= rtol ||D^-1 b||; at a tight rtol it is the dense
    solution; the problem needs several restart cycles."""
    shape = (10, 9, 12)
    img = G.smoothed_threshold(shape, 1.0, 0.55, 5)
    img[0] = G.AIR
    o = make(shape, img, bc_mode=L.BC_ZFACES, press_rtol=rtol)
    o.set_ksp(1)
    o.upload(L.F_SATURATION, np.random.default_rng(5).uniform(0.08, 0.6, size=shape))
    o.phase(L.PH_PROPERTIES)
    t = 77.0
    r, c, v, b = o.pressure_system(t)
    rep = o.phase(L.PH_PRESSURE, t)
    p = o.download(L.F_PRESSURE).ravel()
    A = dense(r, c, v, p.size)
    d = np.diag(A)
    pres = np.linalg.norm((b - A @ p) / d) / np.linalg.norm(b / d)
    assert pres <= rtol * 1.001, pres
    assert abs(rep.press_relres - pres) <= 1e-3 * rtol + 1e-15
    assert rep.press_outer == 2 and rep.press_iters > 30
    if rtol < 1e-10:
        x = np.linalg.solve(A, b)
        np.testing.assert_allclose(p, x, rtol=0, atol=1e-8 * np.abs(x).max())
        # a converged warm start needs no iteration
        rep2 = o.phase(L.PH_PRESSURE, t)
        assert rep2.press_iters == 0


def test_reduced_operator_is_symmetric_positive():
    shape = (6, 6, 6)
    img = G.smoothed_threshold(shape, 1.0, 0.6, 4)
    img[0] = G.AIR
    o = make(shape, img, bc_mode=L.BC_ZFACES)
    o.phase(L.PH_PROPERTIES)
    o.phase(L.PH_PRESSURE, 10.0)
    cls = o.pressure_classes().ravel()
    idx = np.flatnonzero(cls == 1)
    n = img.size
    A = np.zeros((idx.size, idx.size))
    for col, v in enumerate(idx):
        e = np.zeros(n)
        e[v] = 1.0
        A[:, col] = o.apply_pressure(e.reshape(shape)).ravel()[idx]
    np.testing.assert_allclose(A, A.T, rtol=1e-14, atol=0)
    assert np.linalg.eigvalsh(A).min() > 0


# ---------------------------------------------------------------------------
# Flux: linear pressure -> constant gradient; gravity-only flux; clamps
# ---------------------------------------------------------------------------
def test_flux_linear_pressure_and_gravity():
    shape = (6, 5, 4)
    img = np.full(shape, G.PORE, np.uint8)
    o = make(shape, img, max_cap_grad=5.0)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    p = 3.0 * i - 2.0 * j + 0.5 * k
    o.upload(L.F_PRESSURE, p)
    s = np.full(shape, 0.1)
    s[:, :, 2:] = 0.9   # Pc jump along k: gradient clamped to +-5
    o.upload(L.F_SATURATION, s)
    o.phase(L.PH_PROPERTIES)
    o.phase(L.PH_FLUX)
    K = o.download(L.F_PERM)
    T = K / 1e-3
    pc = o.download(L.F_PC)
    qx, qy, qz = o.download(L.F_QX), o.download(L.F_QY), o.download(L.F_QZ)
    np.testing.assert_allclose(qx, -T * (3.0 - 9810.0), rtol=1e-13)
    np.testing.assert_allclose(qy, -T * (-2.0), rtol=1e-13)
    gz = np.gradient(pc, axis=2)    # central inside, one-sided at the faces
    assert np.abs(gz).max() > 5.0
    np.testing.assert_allclose(qz, -T * (0.5 - np.clip(gz, -5.0, 5.0)), rtol=1e-12)


# ---------------------------------------------------------------------------
# Saturation: per-voxel closed form, sub-step count, cap and NaN branch
# ---------------------------------------------------------------------------
def test_saturation_closed_form_and_substeps():
    shape = (8, 3, 3)
    img = np.full(shape, G.PORE, np.uint8)
    img[:, 0, 0] = G.ROCK
    o = make(shape, img, porosity=0.2)
    a = 0.01
    i = np.arange(shape[0], dtype=float)[:, None, None] * np.ones(shape)
    o.upload(L.F_QX, a * i)
    s0 = np.random.default_rng(0).uniform(0.2, 0.9, size=shape)
    o.upload(L.F_SATURATION, s0)
    dt = 3.0
    rep = o.phase(L.PH_SATURATION, 0.0, dt)
    # face averages, zero outside the grid
    qf = a * (np.arange(9) - 0.5)
    qf[0] = 0.0
    qf[-1] = 0.0
    div = np.diff(qf)
    mx = np.abs(div).max()
    assert rep.sat_max_divq == pytest.approx(mx, rel=1e-14)
    nsub = math.ceil(dt / (0.1 * 0.2 / mx))
    assert rep.sat_substeps == nsub
    d = (dt / nsub) / 0.2 * div[:, None, None] * np.ones(shape)
    expect = np.clip(s0 - nsub * d, 0.0, 1.0)   # monotone clamp: N steps = one
    expect[img == G.ROCK] = s0[img == G.ROCK]   # inactive voxels keep s
    np.testing.assert_allclose(o.download(L.F_SATURATION), expect, rtol=0, atol=1e-14 * nsub)


def test_saturation_substep_cap_and_nan():
    shape = (4, 3, 3)
    img = np.full(shape, G.PORE, np.uint8)
    o = make(shape, img)
    q = np.zeros(shape)
    q[1] = 1e6
    o.upload(L.F_QX, q)
    assert o.phase(L.PH_SATURATION, 0.0, 1.0).sat_substeps == 1000
    q[1] = np.nan
    o.upload(L.F_QX, q)
    o.upload(L.F_SATURATION, np.full(shape, 0.5))
    assert o.phase(L.PH_SATURATION, 0.0, 1.0).sat_substeps == 1


# ---------------------------------------------------------------------------
# Surface effects and precipitation: rule tables
# ---------------------------------------------------------------------------
def test_surface_effects_rules():
    shape = (3, 3, 3)
    img = np.full(shape, G.PORE, np.uint8)
    img[0, 1, 1] = G.AIR
    img[2, 2, 2] = G.AIR
    s = np.full(shape, 0.1)
    s[1, 2, 2] = 0.0          # dry pore next to Air: unchanged
    o = make(shape, img)
    o.upload(L.F_SATURATION, s)
    o.phase(L.PH_SURFACE)
    out = o.download(L.F_IMG)
    conc = o.download(L.F_CONC)
    expect = img.copy()
    for (a, b, c) in [(1, 1, 1), (0, 0, 1), (0, 2, 1), (0, 1, 0), (0, 1, 2),
                      (2, 1, 2), (2, 2, 1)]:
        expect[a, b, c] = G.PRECIPITATE
    assert np.array_equal(out, expect)
    assert np.array_equal(conc == 1.0, expect == G.PRECIPITATE)


@pytest.mark.parametrize("instant", [0, 1])
def test_precipitation_rules(instant):
    shape = (2, 2, 3)
    img = np.array([G.PORE] * 6 + [G.ROCK, G.PORE, G.PORE, G.PORE, G.SULPHIDE, G.PORE],
                   np.uint8).reshape(shape)
    conc = np.array([0.5, 0.95, 0.96, 1.0, 2.0, 0.99, 2.0, 0.94, 0.951, 1.5, 3.0, 0.97])
    inv = np.array([0, 0, 0, 0.89, 0.5, 0.9, 0, 0, 0, 0.0, 0, 0.85])
    s = np.full(12, 0.1)
    o = make(shape, img, instant_precip=instant, solid_density=0.5, time_scalar=2.0)
    o.upload(L.F_CONC, conc.reshape(shape))
    o.upload(L.F_INVENTORY, inv.reshape(shape))
    o.upload(L.F_SATURATION, s.reshape(shape))
    o.phase(L.PH_PRECIPITATION)
    lab = o.download(L.F_IMG).ravel()
    c2 = o.download(L.F_CONC).ravel()
    i2 = o.download(L.F_INVENTORY).ravel()
    for v in range(12):
        if img.ravel()[v] == G.PORE and conc[v] > 0.95:
            assert c2[v] == 0.95
            if instant:
                assert lab[v] == G.PRECIPITATE and i2[v] == inv[v]
            else:
                add = (conc[v] - 0.95) * 0.1 * 0.1 / 0.5 * 2.0
                assert i2[v] == pytest.approx(inv[v] + add, rel=1e-15)
                assert lab[v] == (G.PRECIPITATE if inv[v] + add > 0.9 else G.PORE)
        else:
            assert lab[v] == img.ravel()[v] and c2[v] == conc[v] and i2[v] == inv[v]


# ---------------------------------------------------------------------------
# Transport: dense solve, maximum principle, conservation, shared operator
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("theta", [1.0, 0.5])
@pytest.mark.parametrize("ksp", [0, 1])
def test_transport_matches_dense_solve(theta, ksp):
    """ksp 0: BiCGStab; ksp 1: the reference's GMRES(30) + left Jacobi."""
    shape = (5, 6, 4)
    rng = np.random.default_rng(3)
    img = rng.choice([G.AIR, G.PORE, G.PORE, G.ROCK, G.SULPHIDE, G.PRECIPITATE],
                     size=shape).astype(np.uint8)
    o = make(shape, img, theta=theta, adr_rtol=1e-14, alpha_transfer=0.3, Dpore_fac=1.7)
    o.set_ksp(ksp)
    for f in (L.F_QX, L.F_QY, L.F_QZ):
        o.upload(f, rng.uniform(-1, 1, size=shape))
    c0 = rng.uniform(0, 1, size=shape)
    a0 = rng.uniform(0, 1, size=shape)
    o.upload(L.F_CONC, c0)
    o.upload(L.F_ACID, a0)
    dt = 0.4
    r, c, v, b = o.transport_system(dt, L.F_CONC)
    A = dense(r, c, v, img.size)
    xc = np.linalg.solve(A, b)
    r2, c2, v2, b2 = o.transport_system(dt, L.F_ACID)
    A2 = dense(r2, c2, v2, img.size)
    np.testing.assert_array_equal(A, A2)          # one operator for both fields
    xa = np.linalg.solve(A2, b2)
    assert o.phase(L.PH_TRANSPORT, 0.0, dt).status == 0
    np.testing.assert_allclose(o.download(L.F_CONC).ravel(), xc, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(o.download(L.F_ACID).ravel(), xa, rtol=1e-10, atol=1e-12)
    # Precipitate rows keep the old value; Rock/Precipitate flux is zeroed
    prec = img.ravel() == G.PRECIPITATE
    np.testing.assert_array_equal(o.download(L.F_CONC).ravel()[prec], c0.ravel()[prec])
    q = o.download(L.F_QX).ravel()
    assert np.all(q[(img.ravel() == G.ROCK) | prec] == 0.0)


def test_transport_conservation_and_maximum_principle():
    shape = (6, 5, 7)
    img = np.full(shape, G.PORE, np.uint8)
    rng = np.random.default_rng(5)
    c0 = rng.uniform(0.2, 0.8, size=shape)
    o = make(shape, img, adr_rtol=1e-14, Dpore_fac=2.0)
    o.upload(L.F_CONC, c0)
    o.upload(L.F_ACID, c0[::-1].copy())
    o.phase(L.PH_TRANSPORT, 0.0, 2.5)
    c1 = o.download(L.F_CONC)
    assert c1.sum() == pytest.approx(c0.sum(), rel=1e-12)
    assert c1.min() >= c0.min() - 1e-12 and c1.max() <= c0.max() + 1e-12
    assert c1.std() < c0.std()


# ---------------------------------------------------------------------------
# Leaching: single-face closed form, threshold flip, consumption balance
# ---------------------------------------------------------------------------
def test_leaching_single_face_closed_form():
    shape = (3, 3, 3)
    img = np.full(shape, G.PRECIPITATE, np.uint8)
    img[1, 1, 1] = G.SULPHIDE
    img[2, 1, 1] = G.PORE
    o = make(shape, img, D=2.0, kreac=3.0, dx=1.0, time_scalar=1.0)
    acid = np.zeros(shape)
    acid[2, 1, 1] = 0.01
    o.upload(L.F_ACID, acid)
    rep = o.phase(L.PH_LEACHING, 0.0, 0.5)
    k_eff = 1.0 / (1.0 / 2.0 + 1.0 / 3.0)
    loss = k_eff * 0.01 * 0.5
    assert rep.leach_substeps == 1
    assert rep.leach_max_rate == pytest.approx(k_eff * 0.01, rel=1e-15)
    assert o.download(L.F_FRAC)[1, 1, 1] == np.float32(1.0 - loss)
    assert o.download(L.F_ACID)[2, 1, 1] == pytest.approx(0.01 - loss, rel=1e-14)


def test_leaching_threshold_and_balance():
    shape = (6, 6, 6)
    rng = np.random.default_rng(8)
    img = rng.choice([G.PORE, G.ROCK, G.SULPHIDE], size=shape, p=[0.4, 0.3, 0.3]).astype(np.uint8)
    o = make(shape, img, time_scalar=1.0, kreac=0.5)
    frac0 = np.where(img == G.SULPHIDE, rng.uniform(0.3, 1.0, size=shape), 0).astype(np.float32)
    acid0 = rng.uniform(0.0, 0.2, size=shape)
    o.upload(L.F_FRAC, frac0)
    o.upload(L.F_ACID, acid0)
    rep = o.phase(L.PH_LEACHING, 0.0, 0.3)
    frac1, acid1 = o.download(L.F_FRAC), o.download(L.F_ACID)
    assert rep.leached == 0
    # time_scalar 1: acid consumed = solid dissolved (float frac rounding)
    assert (acid0 - acid1).sum() == pytest.approx((frac0 - frac1).astype(float).sum(), rel=1e-5)
    # push one voxel through the threshold
    v = tuple(np.argwhere(img == G.SULPHIDE)[0])
    frac1[v] = 1e-9
    o.upload(L.F_FRAC, frac1)
    acid1[:] = 0.0
    o.upload(L.F_ACID, acid1)
    o.upload(L.F_CONC, np.zeros(shape))
    rep = o.phase(L.PH_LEACHING, 0.0, 0.3)
    assert rep.leached == 1
    assert o.download(L.F_IMG)[v] == G.PORE and o.download(L.F_FRAC)[v] == 0.0
    assert o.download(L.F_CONC)[v] == 0.95 and o.download(L.F_INVENTORY)[v] == 0.0


def test_leaching_substep_count_and_cap():
    shape = (3, 3, 3)
    img = np.full(shape, G.PORE, np.uint8)
    img[1, 1, 1] = G.SULPHIDE
    o = make(shape, img, time_scalar=7.0)
    acid = np.full(shape, 0.02)
    o.upload(L.F_ACID, acid)
    dt = 4.0
    rate = 6 * 0.5 * 0.02
    rep = o.phase(L.PH_LEACHING, 0.0, dt)
    assert rep.leach_substeps == math.ceil(dt / (0.1 / (rate * 7.0)))
    o.upload(L.F_ACID, np.full(shape, 1e3))
    o.upload(L.F_FRAC, np.where(img == G.SULPHIDE, 1.0, 0.0).astype(np.float32))
    assert o.phase(L.PH_LEACHING, 0.0, dt).leach_substeps == 1000


# ---------------------------------------------------------------------------
# Step glue: rollback, initial state, geometry
# ---------------------------------------------------------------------------
def test_oracle_rollback_and_failure():
    img = G.smoothed_threshold((12, 12, 12), 1.5, 0.5, 2)
    o = make(img.shape, img, bc_mode=L.BC_ZFACES)
    o.step(0.0, 1.0)
    before = o.state()
    o.set_params(L.default_params(bc_mode=L.BC_ZFACES, press_maxit=1))
    assert o.step(1.0, 1.0).status == 1
    o.restoreState()
    after = o.state()
    for f in before:
        assert np.array_equal(before[f], after[f])


def test_initial_state_matches_reference_init():
    img = G.reference_channel(0, n=50)
    o = make(img.shape, img)
    assert o.initial_sulphide_count == int((img == G.SULPHIDE).sum()) == 10 * 6 * 6
    s = o.download(L.F_SATURATION)
    assert np.all(s[(img == G.PORE) | (img == G.SULPHIDE)] == 0.1)
    assert np.all(s[(img == G.AIR) | (img == G.ROCK)] == 0.0)
    assert np.all(o.download(L.F_FRAC)[img == G.SULPHIDE] == 1.0)


def test_reference_channel_geometry():
    c1 = G.reference_channel(1)
    assert c1.shape == (50, 50, 50)
    assert (c1 == G.AIR).sum() == 4 * 2500
    assert (c1 == G.PORE).sum() == 46 * 4          # Air overwrites i = 0, 1, 48, 49
    assert (c1 == G.SULPHIDE).sum() == 46 * 36 - 46 * 4


def test_smoothed_threshold_exact_porosity_and_determinism():
    a = G.smoothed_threshold((20, 21, 22), 2.0, 0.3, 7)
    b = G.smoothed_threshold((20, 21, 22), 2.0, 0.3, 7)
    assert np.array_equal(a, b)
    assert (a == G.PORE).sum() == round(0.3 * a.size)
    c = G.multiphase((24, 24, 24), seed=3, fracture_width=0)
    n = c.size
    for lab, fr in ((G.ROCK, 0.30), (G.SULPHIDE, 0.25), (G.PRECIPITATE, 0.15)):
        assert (c == lab).sum() == round(fr * n)
