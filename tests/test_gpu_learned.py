"""Learned preconditioner plumbing on the GPU: the restriction that builds
the network input, and flexible CG reaching the FP64 tolerance with
arbitrary (untrained) weights, falling back to plain PCG if needed."""
import numpy as np
import pytest

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L
from leachingsim_amd import unet_spec as S
from tests.helpers import relerr, small_mixed_geometry

pytestmark = pytest.mark.gpu


def bf16(x):
    from oracle.unet_ref import bf16_round
    return bf16_round(x)


def numpy_pack(r, cls, faces, dims):
    """Restatement of the restriction: sum of the 8 children's residual (ch0,
    before scaling), fraction of solved children, mean +/- faces."""
    ni, nj, nk = r.shape
    out = np.zeros(dims + (16,))
    cx, cy, cz = faces
    cxm = np.zeros_like(cx); cxm[1:] = cx[:-1]
    cym = np.zeros_like(cy); cym[:, 1:] = cy[:, :-1]
    czm = np.zeros_like(cz); czm[:, :, 1:] = cz[:, :, :-1]
    chans = [r, (cls == 1).astype(float), cx, cy, cz, cxm, cym, czm]
    for q, f in enumerate(chans):
        acc = np.zeros(dims)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    sub = f[a::2, b::2, c::2]
                    acc[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += sub
        out[..., q] = acc if q == 0 else acc / 8.0
    m = out[..., 0]
    out[..., 0] = m / np.sqrt((m ** 2).mean())
    return out


def test_learned_pack_matches_restriction():
    shape = (20, 18, 22)
    img = small_mixed_geometry(shape, 21)
    P = L.default_params(bc_mode=L.BC_ZFACES)
    g = L.GpuSimulation(shape, P)
    g.read_labels(img)
    g.init_state()
    g.set_network(np.zeros(S.n_weights(), np.float32))
    g.phase(L.PH_PROPERTIES)
    g.phase(L.PH_PRESSURE, 900.0)
    sd, bh, cl = g.pressure_system()
    rng = np.random.default_rng(0)
    r = np.where(cl == 1, rng.standard_normal(shape), 0.0).astype(np.float32)
    got = g.learned_pack(r)
    # faces: recover the FP32 scaled faces from the unscaled operator is not
    # exposed; compare the channels that do not depend on them exactly and
    # the face channels through their relation to the mask
    dims = g.network_grid()
    assert dims == (16, 16, 16)
    ref = numpy_pack(r.astype(np.float64), cl, [np.zeros(shape)] * 3, dims)
    np.testing.assert_allclose(got[..., 0], bf16(ref[..., 0]), rtol=1e-2, atol=1e-2)
    np.testing.assert_array_equal(got[..., 1], bf16(ref[..., 1]))
    assert np.all(got[..., 8:] == 0)
    faces = got[..., 2:8]
    assert np.all(faces >= 0) and np.all(faces <= 1.0)
    assert np.all(faces[got[..., 1] == 0] == 0)   # no solved child -> no faces


@pytest.mark.parametrize("seed", [1, 2])
def test_learned_solver_reaches_tolerance_and_matches_plain(seed):
    img = G.smoothed_threshold((32, 32, 48), 2.0, 0.4, seed)
    sol = {}
    for solver in (L.SOLVER_PLAIN64, L.SOLVER_LEARNED):
        P = L.default_params(bc_mode=L.BC_ZFACES, solver=solver, press_rtol=1e-12)
        g = L.GpuSimulation(img.shape, P)
        g.read_labels(img)
        g.init_state()
        if solver == L.SOLVER_LEARNED:
            g.set_network(S.init_weights(seed))
        g.phase(L.PH_PROPERTIES)
        rep = g.phase(L.PH_PRESSURE, 1500.0)
        assert rep.status == 0 and rep.press_relres <= 1e-12
        sol[solver] = g.download(L.F_PRESSURE)
        g.close()
    # two different Krylov paths stopped at 1e-12 differ by ~ cond(A) x 1e-12
    assert relerr(sol[L.SOLVER_LEARNED], sol[L.SOLVER_PLAIN64]) <= 1e-7
