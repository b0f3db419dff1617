"""Pins the float64 U-Net reference: central finite differences of the
weight gradient, and the adjoint identity <conv(x), g> = <x, conv^T(g)>."""
import numpy as np

from leachingsim_amd import unet_spec as S
from oracle import unet_ref as U


def test_unet_reference_gradient_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((8, 8, 8, S.CIN))
    w = S.init_weights(1).astype(np.float64)
    w += 0.01 * rng.standard_normal(w.size)
    out, cache = U.forward(x, w)
    gout = rng.standard_normal(out.shape)
    g = U.backward(gout, w, cache)
    eps = 1e-6
    offs = np.cumsum([0] + [a + b for a, b in S.layer_sizes()])
    for lo, hi in zip(offs[:-1], offs[1:]):       # two coordinates per layer
        for i in rng.integers(lo, hi, 2):
            wp, wm = w.copy(), w.copy()
            wp[i] += eps
            wm[i] -= eps
            fd = ((U.forward(x, wp)[0] - U.forward(x, wm)[0]) * gout).sum() / (2 * eps)
            assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, g[i], fd)


def test_conv_adjoint_identity():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((6, 5, 7, 4))
    W = rng.standard_normal((27, 3, 4))
    gy = rng.standard_normal((6, 5, 7, 3))
    y = U.conv(x, W, np.zeros(3))
    gx, gW, _ = U.conv_bwd(x, W, gy)
    assert np.isclose((y * gy).sum(), (x * gx).sum(), rtol=1e-12)
    assert np.isclose((y * gy).sum(), (W * gW).sum(), rtol=1e-12)


def test_pool_up_adjoint():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 8, 3))
    g = rng.standard_normal((2, 3, 4, 3))
    assert np.isclose((U.pool(x) * g).sum(), (x * U.pool_bwd(g)).sum())
    assert np.isclose((U.up(g) * x).sum(), (g * U.up_bwd(x)).sum())


def test_bf16_rounding_ties_to_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float32)
    r = U.bf16_round(v)
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5]


def test_reference_gradient_homogeneity_groups():
    """Zero biases: grad_l . w_l = dL/da of scaling layer l.  Layers on every
    path give L itself; layers in series on one branch share one value."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((8, 8, 8, S.CIN))
    w = S.init_weights(2).astype(np.float64)
    sizes = S.layer_sizes()
    offs = np.cumsum([0] + [a + b for a, b in sizes])
    for li, (nw, nb) in enumerate(sizes):
        w[offs[li] + nw:offs[li + 1]] = 0.0
    out, cache = U.forward(x, w)
    gout = rng.standard_normal(out.shape)
    Lval = float((out * gout).sum())
    g = U.backward(gout, w, cache)
    v = {L["name"]: float(g[offs[li]:offs[li] + sizes[li][0]] @ w[offs[li]:offs[li] + sizes[li][0]])
         for li, L in enumerate(S.LAYERS)}
    for name in ("e0a", "e0b", "d0a", "d0b", "out"):
        assert np.isclose(v[name], Lval, rtol=1e-10), (name, v[name], Lval)
    for group in (("e1a", "e1b", "d1a", "d1b"), ("e2a", "e2b")):
        assert np.allclose([v[n] for n in group], v[group[0]], rtol=1e-10), (group, v)
    # and the branch values are not L (the skip paths are not scaled)
    assert not np.isclose(v["e2a"], Lval, rtol=1e-3)
