"""Independent pins of the transport operator (src/Simulation.cpp:1132-1231).

Each case is a scene small enough to solve by hand: the expected values are
either literals worked out from the reference's row rules or a recursion /
closed form derived from them, never the oracle's own assembly.  Every case
runs on the CPU oracle and, marked ``gpu``, through the C ABI on the device,
so both are held to the same numbers.

The reference's rules being pinned:
  * Pore/Air/Sulphide rows (``set_pore_link``, :1178-1224): for each valid
    neighbour with outward face flux F (+face: q[v], -face: -q[v - stride]),
    D_e = D * Dpore_fac if the neighbour is Pore, else D; F is replaced by 0
    next to Precipitate or Sulphide; cn = -D_e/dx^2, cs = +D_e/dx^2;
    F/dx > 0 adds F/dx to cs, otherwise cn -= F/dx.  A_vn += theta cn,
    A_vv += theta cs, b_v += (1 - theta) * -(cs c_v + cn c_n).
    NOTE the inflow sign: for F < 0 the reference *subtracts* F/dx from cn,
    so an upstream value enters the row with +|F|/dx (:1211-1214), not the
    textbook -|F|/dx.  The pins follow the reference.
  * A Rock neighbour of such a row: A_vv += alpha, A_vn -= alpha, no theta
    (:1187-1194).
  * Rock rows (:1145-1173): 1/dt + alpha per Pore neighbour, -alpha to each
    Pore neighbour; no link to Air, Sulphide, Rock or Precipitate.
  * Precipitate rows (:1132-1140): identity, b = old value.
"""
import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def sim(backend, shape, img, **kw):
    P = L.default_params(adr_rtol=1e-14, **kw)
    if backend == "gpu":
        s = L.GpuSimulation(shape, P)
    else:
        from oracle.oracle import OracleSimulation
        s = OracleSimulation(shape, P)
    s.read_labels(img)
    s.init_state()
    return s


def run_transport(s, dt, conc, qx=None, qy=None, qz=None, acid=None):
    shape = conc.shape
    zero = np.zeros(shape)
    s.upload(L.F_QX, zero if qx is None else qx)
    s.upload(L.F_QY, zero if qy is None else qy)
    s.upload(L.F_QZ, zero if qz is None else qz)
    s.upload(L.F_CONC, conc)
    s.upload(L.F_ACID, zero if acid is None else acid)
    r = s.phase(L.PH_TRANSPORT, 0.0, dt)
    assert r.status == 0
    return s.download(L.F_CONC), s.download(L.F_ACID)


def close(s):
    if hasattr(s, "close"):
        s.close()


# ---------------------------------------------------------------------------
# 1-D pure advection (D = 0) along i with uniform q > 0: each (j, k) column is
# lower bidiagonal, solved by forward substitution.
#   a_ii = 1/dt + theta q/dx  (i < n-1: the +i face exists; outflow)
#        = 1/dt              (i = n-1)
#   a_i,i-1 = +theta q/dx    (i > 0: inflow from i-1, reference sign)
#   rhs_i = c_i/dt - (1-theta) q/dx (c_i [i < n-1] + c_{i-1} [i > 0])
# ---------------------------------------------------------------------------
def upwind_column(c_old, q, dt, dx, theta):
    n = len(c_old)
    c = np.zeros(n)
    a = q / dx
    for i in range(n):
        out = i < n - 1
        inn = i > 0
        diag = 1.0 / dt + (theta * a if out else 0.0)
        rhs = c_old[i] / dt - (1.0 - theta) * ((a * c_old[i] if out else 0.0)
                                               + (a * c_old[i - 1] if inn else 0.0))
        if inn:
            rhs -= theta * a * c[i - 1]
        c[i] = rhs / diag
    return c


@pytest.mark.parametrize("backend", BACKENDS)
def test_advection_column_literal(backend):
    # n = 3, q = 0.5, dt = dx = 1, theta = 1, c_old = (1, 0, 0):
    #   c0 = 1 / 1.5 = 2/3;  c1 = -0.5 c0 / 1.5 = -2/9;  c2 = -0.5 c1 / 1 = 1/9
    # (textbook upwind would give c1 = +2/9)
    shape = (3, 2, 2)
    img = np.full(shape, G.PORE, np.uint8)
    s = sim(backend, shape, img, D=0.0)
    c0 = np.zeros(shape)
    c0[0] = 1.0
    c, _ = run_transport(s, 1.0, c0, qx=np.full(shape, 0.5))
    close(s)
    for j in range(2):
        for k in range(2):
            np.testing.assert_allclose(c[:, j, k], [2 / 3, -2 / 9, 1 / 9], rtol=1e-12)


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_advection_column_recursion(backend, theta):
    shape = (12, 2, 4)
    img = np.full(shape, G.PORE, np.uint8)
    rng = np.random.default_rng(11)
    c_old = rng.uniform(0.0, 1.0, size=shape)
    a_old = rng.uniform(0.0, 1.0, size=shape)
    q, dt, dx = 0.7, 0.8, 1.3
    s = sim(backend, shape, img, D=0.0, theta=theta, dx=dx)
    c, a = run_transport(s, dt, c_old, qx=np.full(shape, q), acid=a_old)
    close(s)
    for j in range(shape[1]):
        for k in range(shape[2]):
            np.testing.assert_allclose(c[:, j, k], upwind_column(c_old[:, j, k], q, dt, dx, theta),
                                       rtol=1e-11, atol=1e-13)
            np.testing.assert_allclose(a[:, j, k], upwind_column(a_old[:, j, k], q, dt, dx, theta),
                                       rtol=1e-11, atol=1e-13)


# ---------------------------------------------------------------------------
# Face rule table, read off the explicit scheme (theta = 0: every non-Rock row
# is diagonal 1/dt, so c_new = c - dt * sum_faces (cs c_v + cn c_n)).
# Scene (2, 2, 2): v = (0,0,0) of type SELF, n = (1,0,0) of type T across
# v's +i face with flux qx[v] = F; v's j and k neighbours are Precipitate
# at c = 0 (each adds cs = D/dx^2 with the barrier).  D = 0.5, Dpore_fac = 3,
# dx = 1, dt = 0.1.
#   c_old = e_n  ->  c_new[v] = -dt * cn
#   c_old = e_v  ->  c_new[v] = 1 - dt * (cs + 2 * 0.5)
# Hand-worked cn / cs (|F| = 0.2):
#   Pore  F>0: D_e 1.5, cn -1.5,          cs 1.5 + 0.2 = 1.7
#   Pore  F<0:          cn -1.5 + 0.2,    cs 1.5
#   Air   F>0: D_e 0.5, cn -0.5,          cs 0.7
#   Air   F<0:          cn -0.3,          cs 0.5
#   Sulphide / Precipitate, either sign: flux barrier, cn -0.5, cs 0.5
# ---------------------------------------------------------------------------
RULES = [  # (neighbour, F, expected c_new[v] for e_n, expected c_new[v] for e_v)
    (G.PORE, +0.2, 0.15, 0.73),
    (G.PORE, -0.2, 0.13, 0.75),
    (G.AIR, +0.2, 0.05, 0.83),
    (G.AIR, -0.2, 0.03, 0.85),
    (G.SULPHIDE, +0.2, 0.05, 0.85),
    (G.SULPHIDE, -0.2, 0.05, 0.85),
    (G.PRECIPITATE, +0.2, 0.05, 0.85),
    (G.PRECIPITATE, -0.2, 0.05, 0.85),
]


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("self_type", [G.PORE, G.AIR, G.SULPHIDE])
def test_face_rule_table(backend, self_type):
    shape = (2, 2, 2)
    for nb_type, F, want_n, want_v in RULES:
        img = np.full(shape, G.PRECIPITATE, np.uint8)
        img[0, 0, 0] = self_type
        img[1, 0, 0] = nb_type
        s = sim(backend, shape, img, D=0.5, Dpore_fac=3.0, theta=0.0)
        qx = np.zeros(shape)
        qx[0, 0, 0] = F
        e_n = np.zeros(shape)
        e_n[1, 0, 0] = 1.0
        e_v = np.zeros(shape)
        e_v[0, 0, 0] = 1.0
        c_n, a_v = run_transport(s, 0.1, e_n, qx=qx, acid=e_v)
        close(s)
        tag = f"self {self_type} neighbour {nb_type} F {F}"
        assert c_n[0, 0, 0] == pytest.approx(want_n, rel=1e-12), tag
        assert a_v[0, 0, 0] == pytest.approx(want_v, rel=1e-12), tag
        # Precipitate rows are identity: their values never move
        prec = img == G.PRECIPITATE
        np.testing.assert_array_equal(c_n[prec], e_n[prec])


# ---------------------------------------------------------------------------
# Warren-Root exchange (D = 0, q = 0, so only the alpha links act).
# Pore-Rock pair: (1/dt + a) c_p - a c_r = c_p0/dt and the mirror row, so
#   c_p + c_r is conserved and (c_p - c_r) = (c_p0 - c_r0) / (1 + 2 a dt),
# for every theta (the Rock link carries no theta).
# Air-Rock and Sulphide-Rock pairs: the active row links to the Rock, the
# Rock row does not link back (it only sees Pore neighbours):
#   c_r = c_r0,   c_x = (c_x0/dt + a c_r0) / (1/dt + a).
# Rock-Rock: no link, both unchanged.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("theta", [1.0, 0.5])
@pytest.mark.parametrize("other", [G.PORE, G.AIR, G.SULPHIDE, G.ROCK])
def test_rock_exchange_closed_form(backend, theta, other):
    shape = (2, 2, 2)
    img = np.full(shape, G.PRECIPITATE, np.uint8)
    img[0, 0, 0] = other
    img[1, 0, 0] = G.ROCK
    a, dt = 0.3, 0.7
    c0 = np.zeros(shape)
    c0[0, 0, 0] = 0.9
    c0[1, 0, 0] = 0.2
    s = sim(backend, shape, img, D=0.0, theta=theta, alpha_transfer=a)
    c, _ = run_transport(s, dt, c0)
    close(s)
    x, r = c[0, 0, 0], c[1, 0, 0]
    if other == G.PORE:
        assert x + r == pytest.approx(1.1, rel=1e-13)
        assert x - r == pytest.approx(0.7 / (1 + 2 * a * dt), rel=1e-12)
    elif other == G.ROCK:
        assert (x, r) == (0.9, 0.2)
    else:
        assert r == 0.2
        assert x == pytest.approx((0.9 / dt + a * 0.2) / (1 / dt + a), rel=1e-12)


@pytest.mark.parametrize("backend", BACKENDS)
def test_rock_with_two_pore_neighbours(backend):
    # Pore - Rock - Pore along i (others Precipitate, D = 0, q = 0).  With
    # u = c_p1 + c_p2, w = c_r: (1/dt + a) u - 2 a w = u0/dt and
    # (1/dt + 2a) w - a u = w0/dt; the total u + w is conserved, so
    # w = (w0/dt + a T) / (1/dt + 3a) with T = u0 + w0, and the Pore
    # difference decays alone: d = d0 / (1 + a dt).
    shape = (3, 2, 2)
    img = np.full(shape, G.PRECIPITATE, np.uint8)
    img[0, 0, 0] = G.PORE
    img[1, 0, 0] = G.ROCK
    img[2, 0, 0] = G.PORE
    a, dt = 0.25, 0.6
    c0 = np.zeros(shape)
    c0[0, 0, 0], c0[1, 0, 0], c0[2, 0, 0] = 1.0, 0.1, 0.4
    s = sim(backend, shape, img, D=0.0, alpha_transfer=a)
    c, _ = run_transport(s, dt, c0)
    close(s)
    T = 1.5
    w = (0.1 / dt + a * T) / (1 / dt + 3 * a)
    assert c[1, 0, 0] == pytest.approx(w, rel=1e-12)
    assert c[0, 0, 0] + c[2, 0, 0] + c[1, 0, 0] == pytest.approx(T, rel=1e-13)
    assert c[0, 0, 0] - c[2, 0, 0] == pytest.approx(0.6 / (1 + a * dt), rel=1e-12)
