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


def small_mixed_geometry(shape, seed, air_slabs=True, sulphide=True, air_axis=0):
    """Pore/Rock field with a Sulphide pocket and one-plane Air slabs on both
    faces of axis `air_axis` (0: the i faces, whose Air voxels carry the
    reference's Dirichlet data; 1: the j faces, which leave the i-face
    inlet/outlet planes of the z-face boundary in contact with Pore, so the
    pressure keeps a drive after the pores next to Air have precipitated)."""
    img = G.smoothed_threshold(shape, 1.5, 0.45, seed)
    if sulphide:
        f = G.smooth_field(shape, 1.5, seed + 7)
        img[(f < np.quantile(f, 0.12)) & (img == G.ROCK)] = G.SULPHIDE
    if air_slabs:
        sl = [slice(None)] * 3
        sl[air_axis] = 0
        img[tuple(sl)] = G.AIR
        sl[air_axis] = -1
        img[tuple(sl)] = G.AIR
    return img
