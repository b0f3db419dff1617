"""Shared helpers for the parity tests (GPU path vs the CPU oracle)."""
from __future__ import annotations

import numpy as np

from leachingsim_amd import geometry as G
from leachingsim_amd import lsgpu as L


def both(shape, params, img, label_bytes=1):
    """A GPU context and an oracle context holding the same initial state."""
    from oracle.oracle import OracleSimulation
    g = L.GpuSimulation(shape, params, label_bytes)
    o = OracleSimulation(shape, params, label_bytes)
    for s in (g, o):
        s.read_labels(img)
        s.init_state()
    return g, o


def copy_state(src, dst, fields=L.STATE_FIELDS):
    for f in fields:
        dst.upload(f, src.download(f))


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b.ravel())
    d = np.linalg.norm((a - b).ravel())
    return d / nb if nb > 0 else d


def small_mixed_geometry(shape, seed, air_slabs=True, sulphide=True):
    """Pore/Rock field with Air slabs on the i faces and a Sulphide pocket."""
    img = G.smoothed_threshold(shape, 1.5, 0.45, seed)
    if sulphide:
        f = G.smooth_field(shape, 1.5, seed + 7)
        img[(f < np.quantile(f, 0.12)) & (img == G.ROCK)] = G.SULPHIDE
    if air_slabs:
        img[0] = G.AIR
        img[-1] = G.AIR
    return img
