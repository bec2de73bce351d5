"""Seeded synthetic models and acquisitions of the benchmark workloads.

Every workload here is generated from a fixed seed (no datasets: nothing is
downloaded).  Shapes follow the project's benchmark configurations:

* ``layered_2d``: 6-12 sinuous interfaces ``z_k(x) = zbar_k + A_k sin(2 pi x /
  lambda_k + phi_k)`` with layer velocities increasing with depth in
  [vmin, vmax], times ``1 + 0.05 n`` where ``n`` is Gaussian-smoothed N(0, 1)
  noise normalised to unit peak;
* ``surface_acquisition``: shots evenly spread in x a few cells below the
  surface, each with an uneven receiver spread (its own count and aperture),
  padded to the largest count with a 0/1 mask for the residual (the
  reference's API takes ``(B, nrec, ndim)`` receivers, so uneven spreads are
  padded, ``core/validation.py:65-71``);
* ``dipping_3d``: planar dipping layers in [vmin, vmax] on a (nz, ny, nx) grid.
"""
from __future__ import annotations

import numpy as np


def ricker(t, f):
    a = (np.pi * f * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _smooth_noise(rng, shape, sigma):
    from scipy.ndimage import gaussian_filter
    n = gaussian_filter(rng.standard_normal(shape), sigma=sigma)
    return n / np.abs(n).max()


def layered_2d(shape, seed=0, vmin=1500.0, vmax=4500.0, noise=0.05, nlayers=None):
    rng = np.random.default_rng(seed)
    nz, nx = shape
    k = int(nlayers if nlayers is not None else rng.integers(6, 13))
    zbar = np.sort(rng.uniform(0.08 * nz, 0.95 * nz, k))
    vel = np.sort(rng.uniform(vmin, vmax, k + 1))
    vel[0] = vmin
    x = np.arange(nx)[None, :]
    z = np.arange(nz)[:, None]
    vp = np.full(shape, vel[0], dtype=np.float64)
    for i in range(k):
        amp = rng.uniform(0.01, 0.05) * nz
        lam = rng.uniform(0.3, 1.5) * nx
        phi = rng.uniform(0, 2 * np.pi)
        iface = zbar[i] + amp * np.sin(2 * np.pi * x / lam + phi)
        vp = np.where(z >= iface, vel[i + 1], vp)
    vp *= 1.0 + noise * _smooth_noise(rng, shape, sigma=4.0)
    return np.clip(vp, vmin, vmax).astype(np.float32)


def smoothed(vp, sigma=12.0):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(np.asarray(vp, dtype=np.float64), sigma=sigma, mode="nearest").astype(np.float32)


def surface_acquisition(shape, nshots, seed=0, nrec_range=(150, 500), depth=2):
    """``(sources (nshots, 2), receivers (nshots, nrec_max, 2), mask (nshots, nrec_max))``
    in physical (x, z) cells."""
    rng = np.random.default_rng(seed + 1000)
    nz, nx = shape
    xs = np.linspace(0.02 * nx, 0.98 * nx, nshots).round().astype(np.int64)
    counts = rng.integers(nrec_range[0], nrec_range[1] + 1, nshots)
    nmax = int(counts.max())
    src = np.stack([xs, np.full(nshots, depth)], axis=1)
    rec = np.zeros((nshots, nmax, 2), dtype=np.int64)
    mask = np.zeros((nshots, nmax), dtype=np.float32)
    for s in range(nshots):
        n = int(counts[s])
        half = rng.uniform(0.2, 0.5) * nx
        lo = int(np.clip(xs[s] - half * rng.uniform(0.3, 1.0), 0, nx - 1))
        hi = int(np.clip(xs[s] + half * rng.uniform(0.3, 1.0), lo + 1, nx - 1))
        rx = np.linspace(lo, hi, n).round().astype(np.int64)
        rec[s, :n, 0] = rx
        rec[s, n:, 0] = rx[-1]
        rec[s, :, 1] = depth
        mask[s, :n] = 1.0
    return src, rec, mask


def patch_acquisition_3d(shape, nshots, patch=64, depth=2):
    """``(sources (nshots, 3), receivers (nshots, patch*patch, 3))`` in
    physical (x, y, z) cells: sources on a regular surface grid, each shot
    recorded by a patch x patch receiver square centred on it (clipped to the
    grid)."""
    nz, ny, nx = shape
    ncol = int(np.ceil(np.sqrt(nshots * nx / ny)))
    nrow = int(np.ceil(nshots / ncol))
    xs = np.linspace(0.2 * nx, 0.8 * nx, ncol).round().astype(np.int64)
    ys = np.linspace(0.2 * ny, 0.8 * ny, max(nrow, 1)).round().astype(np.int64)
    src = np.array([[xs[k % ncol], ys[k // ncol], depth] for k in range(nshots)], dtype=np.int64)
    off = np.arange(patch) - patch // 2
    rec = np.zeros((nshots, patch * patch, 3), dtype=np.int64)
    for s in range(nshots):
        rx = np.clip(src[s, 0] + off, 0, nx - 1)
        ry = np.clip(src[s, 1] + off, 0, ny - 1)
        gy, gx = np.meshgrid(ry, rx, indexing="ij")
        rec[s, :, 0] = gx.ravel()
        rec[s, :, 1] = gy.ravel()
        rec[s, :, 2] = depth
    return src, rec


def dipping_3d(shape, seed=0, vmin=1500.0, vmax=4000.0, nlayers=6):
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z = np.arange(nz)[:, None, None]
    y = np.arange(ny)[None, :, None]
    x = np.arange(nx)[None, None, :]
    depths = np.sort(rng.uniform(0.1 * nz, 0.9 * nz, nlayers))
    vel = np.sort(rng.uniform(vmin, vmax, nlayers + 1))
    vel[0] = vmin
    vp = np.full(shape, vel[0], dtype=np.float32)
    for i in range(nlayers):
        gx, gy = rng.uniform(-0.15, 0.15, 2)
        iface = depths[i] + gx * (x - nx / 2) + gy * (y - ny / 2)
        vp = np.where(z >= iface, np.float32(vel[i + 1]), vp)
    return vp.astype(np.float32)
