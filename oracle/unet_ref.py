"""CPU reference of the 3-level 3D U-Net (forward and backward), float64.

TEST INFRASTRUCTURE ONLY: the values the BF16/MFMA kernels of
leachingsim_amd/csrc/k_unet.hip are checked against (imported by tests/ only).  It has no
counterpart in the reference (the learned preconditioner is new); it
restates the network defined by leachingsim_amd/unet_spec.py, and its
gradients are pinned by central finite differences in the CPU tests.

Conventions (shared with the HIP kernels):
  activations channels-last, shape (ni, nj, nk, C);
  conv weights W[tap, co, ci] with tap = (di+1)*9 + (dj+1)*3 + (dk+1),
  y[v, co] = b[co] + sum_tap sum_ci W[tap, co, ci] x[v + (di,dj,dk), ci],
  zero padding outside the grid;
  pool = 2x2x2 average, up = 2x nearest, concat = [deep, skip].
"""
from __future__ import annotations

import numpy as np

from leachingsim_amd.unet_spec import LAYERS, unpack_weights

OFFS = [(di, dj, dk) for di in (-1, 0, 1) for dj in (-1, 0, 1) for dk in (-1, 0, 1)]


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(x, d):
    """out[v] = x[v + d] with zeros outside (x: (ni, nj, nk, C))."""
    out = np.zeros_like(x)
    src = [slice(max(0, dd), x.shape[a] + min(0, dd)) for a, dd in enumerate(d)]
    dst = [slice(max(0, -dd), x.shape[a] + min(0, -dd)) for a, dd in enumerate(d)]
    out[tuple(dst)] = x[tuple(src)]
    return out


def conv(x, W, b):
    y = np.broadcast_to(b, x.shape[:3] + (W.shape[1],)).astype(np.float64).copy()
    for t, d in enumerate(OFFS):
        y += shift(x, d) @ W[t].T
    return y


def conv_bwd(x, W, gy):
    """gradients of conv: (gx, gW, gb)."""
    gx = np.zeros_like(x)
    gW = np.zeros_like(W)
    for t, d in enumerate(OFFS):
        xs = shift(x, d)
        gW[t] = gy.reshape(-1, gy.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
        gx += shift(gy @ W[t], tuple(-e for e in d))
    gb = gy.sum(axis=(0, 1, 2))
    return gx, gW, gb


def pool(x):
    s = x.shape
    return x.reshape(s[0] // 2, 2, s[1] // 2, 2, s[2] // 2, 2, s[3]).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up_bwd(g):
    s = g.shape
    return g.reshape(s[0] // 2, 2, s[1] // 2, 2, s[2] // 2, 2, s[3]).sum(axis=(1, 3, 5))


def forward(x, flat_w, bf16=False):
    """x: (ni, nj, nk, CIN) float; returns (out (ni,nj,nk), cache).
    bf16=True rounds the input, the weights and every stored activation to
    bfloat16 where the device kernels do (accumulation stays float64)."""
    R = bf16_round if bf16 else (lambda a: a)
    flat_w = np.asarray(flat_w, np.float64)
    if bf16:
        P0 = unpack_weights(flat_w)
        parts = []
        for L in LAYERS:
            W, b = P0[L["name"]]
            parts += [R(W).ravel(), b]
        flat_w = np.concatenate(parts)
    P = unpack_weights(flat_w)
    x = R(x)
    cache = {"x": x}
    h = x
    acts = {}
    for L in LAYERS:
        name = L["name"]
        if L["src"] == "pool":
            h = R(pool(acts[L["from"]]))
        elif L["src"] == "cat":
            h = np.concatenate([up(acts[L["deep"]]), acts[L["skip"]]], axis=-1)
        elif L["src"] == "prev":
            h = acts[L["from"]] if L["from"] else h
        cache["in_" + name] = h
        W, b = P[name]
        y = conv(h, W, b)
        if L["relu"]:
            y = np.maximum(y, 0.0)
            y = R(y)
        acts[name] = y
        h = y
    cache["acts"] = acts
    return acts[LAYERS[-1]["name"]][..., 0], cache


def backward(gout, flat_w, cache, bf16=False):
    """gout: (ni, nj, nk) gradient of the output channel; returns the flat
    weight gradient (same layout as the weights).  bf16=True rounds each
    layer's output gradient to bfloat16 before it enters the convolution
    gradients, as the device's MFMA operands are."""
    R = bf16_round if bf16 else (lambda a: a)
    P = unpack_weights(np.asarray(flat_w, np.float64))
    acts = cache["acts"]
    grads = {}
    gact = {L["name"]: None for L in LAYERS}
    last = LAYERS[-1]
    g = np.zeros(acts[last["name"]].shape)
    g[..., 0] = gout
    gact[last["name"]] = g
    for L in reversed(LAYERS):
        name = L["name"]
        gy = gact[name]
        if gy is None:
            gy = np.zeros_like(acts[name])
        if L["relu"]:
            gy = gy * (acts[name] > 0)
        W, b = P[name]
        gx, gW, _ = conv_bwd(cache["in_" + name], R(W) if bf16 else W, R(gy))
        gb = gy.sum(axis=(0, 1, 2))
        grads[name] = (gW, gb)

        def acc(key, val):
            gact[key] = val if gact[key] is None else gact[key] + val

        if L["src"] == "pool":
            acc(L["from"], pool_bwd(gx))
        elif L["src"] == "cat":
            cd = acts[L["deep"]].shape[-1]
            acc(L["deep"], up_bwd(gx[..., :cd]))
            acc(L["skip"], gx[..., cd:])
        elif L["src"] == "prev" and L["from"]:
            acc(L["from"], gx)
    parts = []
    for L in LAYERS:
        gW, gb = grads[L["name"]]
        parts += [gW.ravel(), gb.ravel()]
    return np.concatenate(parts)
