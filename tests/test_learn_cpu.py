"""Host-side pieces of the learned preconditioner's training: the energy
loss gradient (central differences) and the adjointness of P and P^T."""
import numpy as np

from leachingsim_amd import learn


def setup(seed=0, n=8):
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    mask = (rng.random(shape) < 0.7).astype(float)
    faces = [0.15 * rng.random(shape) * mask for _ in range(3)]
    faces[0][:-1] *= mask[1:]
    faces[1][:, :-1] *= mask[:, 1:]
    faces[2][:, :, :-1] *= mask[:, :, 1:]
    faces[0][-1] = 0
    faces[1][:, -1] = 0
    faces[2][:, :, -1] = 0
    e = rng.standard_normal(shape) * mask
    r = learn.apply_op(e, faces) * mask
    return rng, mask, faces, e, r


def test_operator_symmetric():
    rng, mask, faces, e, r = setup()
    a = rng.standard_normal(mask.shape) * mask
    b = rng.standard_normal(mask.shape) * mask
    assert np.isclose((a * learn.apply_op(b, faces)).sum(), (b * learn.apply_op(a, faces)).sum())


def test_prolong_restrict_adjoint():
    rng, mask, faces, e, r = setup()
    c = rng.standard_normal((16, 16, 16))
    f = rng.standard_normal(mask.shape)
    assert np.isclose((learn.prolong(c, mask) * f).sum(), (c * learn.restrict_sum(f * mask, c.shape)).sum())


def test_energy_loss_gradient_and_bounds():
    rng, mask, faces, e, r = setup(1)
    out = rng.standard_normal((16, 16, 16))
    loss, g = learn.energy_loss_grad(out, r, e, faces, mask)
    assert 0.0 <= loss <= 1.0
    eps = 1e-6
    for idx in [(0, 0, 0), (1, 2, 3), (3, 3, 1), (2, 0, 1)]:
        o1, o2 = out.copy(), out.copy()
        o1[idx] += eps
        o2[idx] -= eps
        fd = (learn.energy_loss_grad(o1, r, e, faces, mask)[0] -
              learn.energy_loss_grad(o2, r, e, faces, mask)[0]) / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-6 * max(1e-3, abs(fd)), (idx, fd, g[idx])
    # the exact error restricted as coarse correction captures a positive share
    loss_e, _ = learn.energy_loss_grad(learn.restrict_sum(e, (16, 16, 16)) / 8, r, e, faces, mask)
    assert loss_e < 1.0


def test_galerkin_operator_is_the_energy_on_the_network_grid():
    """out.A_c out == u.A^u and out.P^T r == u.r for u = P out, on an even and
    an odd grid (the coarse grid padded past the fine one)."""
    for n in (8, 7):
        rng, mask, faces, e, r = setup(2, n)
        dims = (8, 8, 8)
        ac = learn.galerkin_operator(faces, mask, dims)
        out = rng.standard_normal(dims)
        u = learn.prolong(out, mask)
        assert np.isclose((out * learn.coarse_apply(ac, out)).sum(),
                          (u * learn.apply_op(u, faces)).sum(), rtol=1e-12)
        rc = learn.restrict_sum(r * mask, dims)
        assert np.isclose((out * rc).sum(), (u * r).sum(), rtol=1e-12)
        # the coarse operator is symmetric
        a = rng.standard_normal(dims)
        assert np.isclose((a * learn.coarse_apply(ac, out)).sum(),
                          (out * learn.coarse_apply(ac, a)).sum(), rtol=1e-12)


def test_coarse_loss_matches_the_fine_loss():
    rng, mask, faces, e, r = setup(3)
    dims = (8, 8, 8)
    out = rng.standard_normal(dims)
    ac = learn.galerkin_operator(faces, mask, dims)
    rc = learn.restrict_sum(r * mask, dims)
    eae = float((e * r).sum())
    lf, gf = learn.energy_loss_grad(out, r, e, faces, mask)
    lc, gc = learn.coarse_loss_grad(out, rc, eae, ac)
    assert np.isclose(lf, lc, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gc, gf, rtol=1e-5, atol=1e-7 * np.abs(gf).max())
