"""The 3-level 3D U-Net of the learned pressure preconditioner / warm start.

Levels 16/32/64 channels, 3x3x3 convolutions with zero padding, ReLU,
2x2x2 average pooling down, 2x nearest upsampling up, skip concatenation
[deep, skip].  Activations are BF16 channels-last on the device, FP32
accumulation on MFMA.  The final layer has 16 output channels of which only
channel 0 is used (the others have zero weights and receive zero gradient),
so every layer maps onto 16-wide MFMA tiles.

Input channels (CIN = 16, zero padded): 0 the normalised field (residual for
the preconditioner, previous pressure for the warm start), 1 the solved-row
mask, 2-4 the scaled +face conductances (i, j, k), 5-7 the scaled -face
conductances, 8 the ratio of boundary pressures (warm start only), 9-15 zero.

Flat weight layout (FP32): for each layer in order, W[27][Cout][Cin] then
b[Cout].
"""
from __future__ import annotations

import numpy as np

CIN = 16

# src: "input" (network input), "prev" (output of layer `from`), "pool"
# (2x2x2 average of layer `from`), "cat" ([up(deep), skip])
LAYERS = [
    dict(name="e0a", cin=CIN, cout=16, relu=True, src="input", level=0),
    dict(name="e0b", cin=16, cout=16, relu=True, src="prev", level=0, **{"from": "e0a"}),
    dict(name="e1a", cin=16, cout=32, relu=True, src="pool", level=1, **{"from": "e0b"}),
    dict(name="e1b", cin=32, cout=32, relu=True, src="prev", level=1, **{"from": "e1a"}),
    dict(name="e2a", cin=32, cout=64, relu=True, src="pool", level=2, **{"from": "e1b"}),
    dict(name="e2b", cin=64, cout=64, relu=True, src="prev", level=2, **{"from": "e2a"}),
    dict(name="d1a", cin=96, cout=32, relu=True, src="cat", level=1, deep="e2b", skip="e1b"),
    dict(name="d1b", cin=32, cout=32, relu=True, src="prev", level=1, **{"from": "d1a"}),
    dict(name="d0a", cin=48, cout=16, relu=True, src="cat", level=0, deep="d1b", skip="e0b"),
    dict(name="d0b", cin=16, cout=16, relu=True, src="prev", level=0, **{"from": "d0a"}),
    dict(name="out", cin=16, cout=16, relu=False, src="prev", level=0, **{"from": "d0b"}),
]
for _L in LAYERS:
    _L.setdefault("from", None)


def layer_sizes():
    return [(27 * L["cout"] * L["cin"], L["cout"]) for L in LAYERS]


def n_weights() -> int:
    return sum(w + b for w, b in layer_sizes())


def unpack_weights(flat):
    out = {}
    o = 0
    for L, (nw, nb) in zip(LAYERS, layer_sizes()):
        W = flat[o:o + nw].reshape(27, L["cout"], L["cin"])
        o += nw
        b = flat[o:o + nb]
        o += nb
        out[L["name"]] = (W, b)
    return out


def init_weights(seed: int = 0) -> np.ndarray:
    """He-normal initialisation; output layer channels 1..15 zero."""
    rng = np.random.default_rng(seed)
    parts = []
    for L in LAYERS:
        fan_in = 27 * L["cin"]
        W = rng.standard_normal((27, L["cout"], L["cin"])) * np.sqrt(2.0 / fan_in)
        b = np.zeros(L["cout"])
        if L["name"] == "out":
            W *= 0.1
            W[:, 1:, :] = 0.0
        parts += [W.ravel(), b]
    return np.concatenate(parts).astype(np.float32)
