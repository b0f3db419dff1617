"""U-Net kernels (BF16 MFMA, FP32 accumulation) vs the float64 reference.

The reference rounds input, weights and stored activations to bfloat16 at
the same points as the device; the remaining differences are FP32 vs FP64
accumulation and BF16 rounding of the gradients fed to the MFMA, hence the
relative L2 tolerances written below.
"""
import numpy as np
import pytest

from leachingsim_amd import unet_spec as S

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm((np.asarray(a, float) - b).ravel()) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def setup(seed, shape=(16, 16, 32), batch=2):
    from leachingsim_amd.unet import Net
    rng = np.random.default_rng(seed)
    w = S.init_weights(seed).astype(np.float64)
    w += 0.02 * rng.standard_normal(w.size) * (w != 0)
    x = rng.standard_normal((batch,) + shape + (S.CIN,)).astype(np.float32)
    net = Net(shape, batch=batch, trainable=True, weights=w.astype(np.float32))
    return net, w.astype(np.float32).astype(np.float64), x, rng


def test_unet_forward_matches_reference():
    from oracle import unet_ref as U
    net, w, x, _ = setup(1)
    out = net.forward(x)
    for b in range(x.shape[0]):
        ref, _ = U.forward(x[b].astype(np.float64), w, bf16=True)
        assert relerr(out[b], ref) <= 2e-2


def test_unet_backward_matches_reference():
    from oracle import unet_ref as U
    net, w, x, rng = setup(2)
    net.forward(x)
    gout = rng.standard_normal((x.shape[0],) + x.shape[1:4]).astype(np.float32)
    g = net.backward(gout)
    gref = np.zeros_like(w)
    for b in range(x.shape[0]):
        _, cache = U.forward(x[b].astype(np.float64), w, bf16=True)
        gref += U.backward(gout[b].astype(np.float64), w, cache, bf16=True)
    parts = S.unpack_weights(g.astype(np.float64))
    pref = S.unpack_weights(gref)
    errs = {}
    for L in S.LAYERS:
        gw, gb = parts[L["name"]]
        rw, rb = pref[L["name"]]
        if L["name"] == "out":
            gw, rw, gb, rb = gw[:, :1], rw[:, :1], gb[:1], rb[:1]
        cos = float((gw * rw).sum() / (np.linalg.norm(gw) * np.linalg.norm(rw)))
        errs[L["name"]] = (relerr(gw, rw), relerr(gb, rb), cos)
    # BF16 activations and gradients (8-bit significands) through up to 11
    # layers, with cancellation in the voxel sums: 15 % relative L2 and a
    # cosine of at least 0.99 per layer (an indexing error in a tap or a
    # channel block breaks the cosine)
    assert all(a <= 0.15 and b <= 0.15 and c >= 0.99 for a, b, c in errs.values()), errs


def test_unet_gradient_euler_identity():
    """With zero biases every ReLU conv layer is positively homogeneous, so
    d/da L(a w_l)|_{a=1} = grad_l . w_l.  Scaling a layer scales everything
    downstream of it up to the next concatenation, hence:
      - layers every path runs through (e0a e0b d0a d0b out): grad.w = L;
      - layers in series on one branch share one value: {e1a e1b d1a d1b}
        (they scale the whole level-1 branch) and {e2a e2b}.
    Checked on the device's own forward and backward."""
    net, w, x, rng = setup(4, shape=(16, 16, 16), batch=1)
    sizes = S.layer_sizes()
    offs = np.cumsum([0] + [a + b for a, b in sizes])
    for li, (nw, nb) in enumerate(sizes):          # biases zero
        w[offs[li] + nw:offs[li + 1]] = 0.0
    net.set_weights(w.astype(np.float32))
    out = net.forward(x)
    gout = rng.standard_normal(out.shape).astype(np.float32)
    Lval = float((out.astype(np.float64) * gout).sum())
    g = net.backward(gout).astype(np.float64)
    vals = {L["name"]: float(g[offs[li]:offs[li] + sizes[li][0]] @ w[offs[li]:offs[li] + sizes[li][0]])
            for li, L in enumerate(S.LAYERS)}
    # exact in real arithmetic; the BF16 rounding of activations and of the
    # gradients fed to the MFMA adds a few % (the emulated BF16 reference
    # shows up to 8 % of |L| between members of the level-1 group)
    for name in ("e0a", "e0b", "d0a", "d0b", "out"):
        assert abs(vals[name] - Lval) <= 0.10 * abs(Lval), (name, Lval, vals)
    for group in (("e1a", "e1b", "d1a", "d1b"), ("e2a", "e2b")):
        v = np.array([vals[n] for n in group])
        scale = max(abs(v.mean()), abs(Lval))
        assert v.max() - v.min() <= 0.15 * scale, (group, Lval, vals)


@pytest.mark.parametrize("shape,batch", [((16, 16, 32), 2), ((32, 48, 80), 1), ((16, 32, 64), 3)])
def test_tiled_conv_matches_direct(shape, batch):
    """The LDS-tiled convolution kernels (every layer shape: plain, with the
    upsampled concatenation, and backward-data) issue the MFMA sequence of
    the general kernel, so forward output and weight gradient agree bit for
    bit.  nk = 80 leaves a partial k-tile at every level."""
    net, w, x, rng = setup(5, shape=shape, batch=batch)
    gout = rng.standard_normal((batch,) + shape).astype(np.float32)
    res = {}
    for tiled in (True, False):
        net.set_conv(tiled)
        out = net.forward(x)
        res[tiled] = (out, net.backward(gout))
    assert np.array_equal(res[True][0], res[False][0])
    assert np.array_equal(res[True][1], res[False][1])
    assert np.abs(res[True][0]).max() > 0
