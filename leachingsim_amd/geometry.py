"""Seeded synthetic geometries and the .raw label format.

Shapes are (ni, nj, nk) = the reference's internal (i, j, k) = CLI
(zext, yext, xext) (src/main.cpp:94); arrays are C-ordered with k fastest,
which is the .raw byte order (src/MPIRawLoader.h:64-80).

Generators:
  smoothed_threshold   Gaussian white noise, periodic Gaussian smoothing
                       (FFT), thresholded at the exact quantile so the pore
                       count is round(porosity * N).
  multiphase           pore + 3 mineral phases from independent fields, plus
                       an optional planar fracture through the centre.
  reference_channel    the 50^3 test geometries of generate_input.py
                       (closed-form slicing, cases 0 and 1).
  fractured_binary     smoothed_threshold plus two orthogonal planar
                       fractures through the centre and Air face slabs.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

AIR, PORE, ROCK, SULPHIDE, PRECIPITATE = 1, 2, 3, 4, 5


def smooth_field(shape, sigma: float, seed: int, dtype=np.float32) -> np.ndarray:
    """White noise smoothed by a periodic Gaussian of std `sigma` voxels."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape, dtype=np.float32)
    if sigma <= 0:
        return noise.astype(dtype, copy=False)
    F = np.fft.rfftn(noise)
    del noise
    for ax, n in enumerate(shape):
        f = np.fft.rfftfreq(n) if ax == len(shape) - 1 else np.fft.fftfreq(n)
        g = np.exp(-2.0 * (np.pi * sigma * f) ** 2).astype(np.float32)
        bshape = [1] * len(shape)
        bshape[ax] = g.size
        F *= g.reshape(bshape)
    out = np.fft.irfftn(F, s=shape, axes=range(len(shape)))
    return out.astype(dtype, copy=False)


def _lowest(field: np.ndarray, count: int, allowed: np.ndarray | None = None) -> np.ndarray:
    """Boolean mask of the `count` lowest values of `field` (among `allowed`)."""
    flat = field.ravel()
    if allowed is not None:
        idx = np.flatnonzero(allowed.ravel())
        vals = flat[idx]
        sel = idx[np.argpartition(vals, count - 1)[:count]] if count > 0 else idx[:0]
    else:
        sel = np.argpartition(flat, count - 1)[:count] if count > 0 else np.empty(0, np.int64)
    m = np.zeros(flat.size, dtype=bool)
    m[sel] = True
    return m.reshape(field.shape)


def smoothed_threshold(shape, sigma: float, porosity: float, seed: int,
                       dtype=np.uint8) -> np.ndarray:
    """Binary Pore (2) / Rock (3) image with exactly round(porosity*N) pores."""
    f = smooth_field(shape, sigma, seed)
    n = int(round(porosity * f.size))
    img = np.full(shape, ROCK, dtype=dtype)
    img[_lowest(f, n)] = PORE
    return img


def planar_fracture(shape, width: float, normal, center=None) -> np.ndarray:
    """Mask of voxels whose centre lies within width/2 of a plane."""
    n = np.asarray(normal, dtype=np.float64)
    n = n / np.linalg.norm(n)
    c = (np.asarray(shape, dtype=np.float64) - 1.0) / 2.0 if center is None else np.asarray(center)
    ii = (np.arange(shape[0]) - c[0]) * n[0]
    jj = (np.arange(shape[1]) - c[1]) * n[1]
    kk = (np.arange(shape[2]) - c[2]) * n[2]
    d = ii[:, None, None] + jj[None, :, None] + kk[None, None, :]
    return np.abs(d) < width / 2.0


def random_normal(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(3)
    return v / np.linalg.norm(v)


def multiphase(shape, fractions=(0.30, 0.25, 0.15), labels=(ROCK, SULPHIDE, PRECIPITATE),
               sigma: float = 3.0, seed: int = 0, fracture_width: float = 4.0,
               air_j_planes: int = 0, dtype=np.uint16) -> np.ndarray:
    """Pore + mineral phases, each placed from its own seeded smoothed field.

    Phase q takes the round(fraction_q * N) lowest voxels of its field among
    the voxels not yet assigned; the rest is Pore.  A planar fracture of
    `fracture_width` voxels through the centre at a seeded orientation is set
    to Pore; `air_j_planes` planes at each j face are set to Air.
    """
    N = int(np.prod(shape))
    img = np.full(shape, PORE, dtype=dtype)
    free = np.ones(shape, dtype=bool)
    for q, (fr, lab) in enumerate(zip(fractions, labels)):
        f = smooth_field(shape, sigma, seed * 101 + q + 1)
        m = _lowest(f, int(round(fr * N)), allowed=free)
        img[m] = lab
        free &= ~m
        del f
    if fracture_width > 0:
        img[planar_fracture(shape, fracture_width, random_normal(seed * 101 + 99))] = PORE
    if air_j_planes > 0:
        img[:, :air_j_planes, :] = AIR
        img[:, -air_j_planes:, :] = AIR
    return img


def fractured_binary(shape, sigma: float = 3.0, porosity: float = 0.25, seed: int = 1,
                     fracture_width: float = 6.0, air_j_planes: int = 2,
                     dtype=np.uint8) -> np.ndarray:
    """Pore / Rock from smoothed_threshold, then two orthogonal planar
    fractures of `fracture_width` voxels through the centre set to Pore (the
    first at a seeded orientation, the second with a seeded normal in the
    plane of the first) and `air_j_planes` planes of Air at each j face, the
    surface that drives evaporation and precipitation."""
    img = smoothed_threshold(shape, sigma, porosity, seed, dtype=dtype)
    n1 = random_normal(seed * 101 + 98)
    v = random_normal(seed * 101 + 97)
    n2 = v - (v @ n1) * n1
    n2 /= np.linalg.norm(n2)
    for nrm in (n1, n2):
        img[planar_fracture(shape, fracture_width, nrm)] = PORE
    if air_j_planes > 0:
        img[:, :air_j_planes, :] = AIR
        img[:, -air_j_planes:, :] = AIR
    return img


def reference_channel(case: int = 1, n: int = 50, dtype=np.uint8) -> np.ndarray:
    """The synthetic inputs of generate_input.py (closed form).

    case 0: Rock block with a 10x10 Pore channel, a Sulphide block inside it,
            2-plane Air slabs at both i faces (generate_input.py:29-50).
    case 1: Rock with a 6x6 Sulphide vein, a 2x2 Pore channel through it,
            the same Air slabs (generate_input.py:52-70).
    """
    img = np.full((n, n, n), ROCK, dtype=dtype)
    if case == 0:
        img[:, 20:30, 20:30] = PORE
        img[20:30, 22:28, 22:28] = SULPHIDE
        img[0:2] = AIR
        img[-2:] = AIR
    else:
        img[2:48, 22:28, 22:28] = SULPHIDE
        img[:, 24:26, 24:26] = PORE
        img[0:2] = AIR
        img[48:50] = AIR
    return img


# Named workloads
WORKLOADS = ("smooth64", "smooth338", "fractured512", "fractured1024x768")


def named_geometry(name: str, seed: int = 1, size: int | None = None) -> np.ndarray:
    """smooth64: 64^3 uint8, sigma 2, porosity 0.40 (the reference's
    CPU-runnable case); smooth338: 338x338x283 (x, y, z) -> (ni, nj, nk) =
    (283, 338, 338) uint8, sigma 3, porosity 0.30 (README scale);
    fractured512: 512^3 (or size^3) uint16 labels, pore + three minerals
    (0.30 / 0.25 / 0.15), a 4-voxel planar fracture, Air on the j faces;
    fractured1024x768: 1024x1024x768 (x, y, z) -> (ni, nj, nk) =
    (768, 1024, 1024) uint8 (or (3 size / 4, size, size)), sigma 3, porosity
    0.25, two orthogonal 6-voxel fractures, Air on the j faces."""
    if name == "smooth64":
        return smoothed_threshold((64, 64, 64), 2.0, 0.40, seed)
    if name == "smooth338":
        return smoothed_threshold((283, 338, 338), 3.0, 0.30, seed)
    if name == "fractured512":
        n = size or 512
        return multiphase((n, n, n), sigma=3.0, seed=seed, fracture_width=4.0, air_j_planes=2)
    if name == "fractured1024x768":
        n = size or 1024
        return fractured_binary((3 * n // 4, n, n), sigma=3.0, porosity=0.25, seed=seed,
                                fracture_width=6.0, air_j_planes=2)
    raise KeyError(name)


def read_raw(path, shape, dtype=np.uint8, header: int = 0) -> np.ndarray:
    """Headerless (or header-skipped) .raw reader (src/MPIRawLoader.h:48-82)."""
    n = int(np.prod(shape))
    data = np.fromfile(path, dtype=dtype, count=n, offset=header)
    if data.size != n:
        raise ValueError(f"{path}: expected {n} voxels, found {data.size}")
    return data.reshape(shape)


def write_raw(path, img: np.ndarray) -> None:
    np.ascontiguousarray(img).tofile(Path(path))
