"""Mixed-precision CG on the red-black Schur complement: the same solution as
plain FP64 PCG (to the solve tolerance) in about half the iterations of the
inner CG on the operator itself.  The factor is fixed by the mathematics:
the scaled operator is I - N with N 2-cyclic (the 7-point graph is
bipartite), so S = I - N_br N_rb has cond(S) ~ cond(A^) / 4; a numpy
emulation on the oracle's operator of a 96^3 workload gave 1016 -> 508."""
import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L
from tests.helpers import relerr, small_mixed_geometry

pytestmark = pytest.mark.gpu


def solve(img, t, **kw):
    P = L.default_params(bc_mode=L.BC_ZFACES, press_rtol=1e-12, **kw)
    g = L.GpuSimulation(img.shape, P)
    g.read_labels(img)
    g.init_state()
    g.phase(L.PH_PROPERTIES)
    rep = g.phase(L.PH_PRESSURE, t)
    p = g.download(L.F_PRESSURE)
    g.close()
    return rep, p


@pytest.mark.parametrize("seed", [1, 2])
def test_schur_matches_plain_in_fewer_iterations(seed):
    img = G.smoothed_threshold((40, 36, 48), 2.0, 0.4, seed)     # nk % 4 == 0: vector path
    r64, p64 = solve(img, 1500.0, solver=L.SOLVER_PLAIN64)
    rs, ps = solve(img, 1500.0, solver=L.SOLVER_MIXED, mixed_schur=1)
    rm, pm = solve(img, 1500.0, solver=L.SOLVER_MIXED, mixed_schur=0)
    for r in (r64, rs, rm):
        assert r.status == 0 and r.press_relres <= 1e-12
    assert relerr(ps, p64) <= 1e-7 and relerr(pm, p64) <= 1e-7
    print(f"iterations: plain FP64 {r64.press_iters}, mixed {rm.press_iters}, "
          f"mixed Schur {rs.press_iters}")
    assert rs.press_iters <= 0.7 * rm.press_iters


def test_schur_coupled_steps_match_oracle():
    from tests.helpers import both
    img = small_mixed_geometry((24, 20, 28), 9)
    P = L.default_params(bc_mode=L.BC_ZFACES, solver=L.SOLVER_MIXED, mixed_schur=1,
                         press_rtol=1e-12)
    g, o = both(img.shape, P, img)
    t = 0.0
    for n in range(4):
        rg = g.step(t, 30.0)
        ro = o.step(t, 30.0)
        assert rg.status == ro.status == 0
        assert np.array_equal(g.download(L.F_IMG), o.download(L.F_IMG)), n
        t += 30.0
    assert relerr(g.download(L.F_PRESSURE), o.download(L.F_PRESSURE)) <= 1e-7
    assert relerr(g.download(L.F_CONC), o.download(L.F_CONC)) <= 1e-7
