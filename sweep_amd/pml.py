"""CPML damping profiles on the runtime grid (host side, numpy float32).

The kernels read these 1-D profiles per axis; they must be the reference's
values bit for bit (source/receiver maps and CPML masks are a bit-exact
contract), so every numpy operation below is the one the reference applies,
in the same order and the same dtype:

* :func:`ramp_profiles` follows ``equations/pml.py:216-269`` (``setup_pml``):
  ``sigma0 = -(1+n) v_max ln(R) / (2 L)``, ``alpha0 = pi f_pml``,
  ``a = exp(-(sigma+alpha) dt)`` zeroed where the ramp fraction is 0,
  ``b = sigma/(sigma+alpha) (a-1)``;
* :func:`cpml_profiles_r` follows ``equations/pml.py:157-214``
  (``set_cpml_profiles_r``, the second-order ``cpmlr`` form) and appends
  ``db = np.gradient(b, h)``, which is non-zero one cell past each band;
* :func:`cpml_profiles_s` follows ``equations/pml.py:81-154`` (the staggered
  ``cpmls`` form with half-cell profiles ``ah, bh`` at ``start=0.5``).

The bands start ``fd_pad + width`` cells in from each runtime-grid edge
(``pml_start``), so with ``fd_pad = M`` the ramp covers the PML pad exactly.
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np


def ramp_profiles(width: Sequence[int], start_cells: Sequence[float], max_pml: float,
                  dt: float, n: int, max_vel: float, dtype=np.float32, pml_freq: float = 25.0,
                  start: float = 0.0, r_val: float = 0.001, n_power: int = 2, eps: float = 1e-9):
    """``(a, b)`` of length ``n`` for one axis with low/high ramp ``width``."""
    alpha0 = math.pi * pml_freq
    if max_pml == 0:
        return np.zeros(n, dtype=dtype), np.zeros(n, dtype=dtype)
    sigma0 = -(1 + n_power) * max_vel * math.log(r_val) / (2 * max_pml)
    x = np.arange(n, dtype=dtype) + dtype(start)
    lo = np.zeros_like(x) if width[0] == 0 else (dtype(start_cells[0]) - x) / dtype(width[0])
    hi = np.zeros_like(x) if width[1] == 0 else (x - dtype(start_cells[1])) / dtype(width[1])
    frac = np.clip(np.maximum(lo, hi), 0.0, 1.0)
    sigma = dtype(sigma0) * (frac ** n_power)
    alpha = dtype(alpha0) * (dtype(1.0) - frac)
    total = sigma + alpha
    total = np.where(np.abs(total) < eps, dtype(eps), total)
    a = np.exp(-total * abs(dt)).astype(dtype, copy=False)
    b = (sigma / total) * (a - dtype(1.0))
    a = a.copy()
    a[frac == 0] = dtype(0.0)
    return a, b.astype(dtype, copy=False)


def _band_starts(width, fd_pad, shape, dim, as_float):
    lo = fd_pad[2 * dim] + width[2 * dim]
    hi = shape[dim] - 1 - fd_pad[2 * dim + 1] - width[2 * dim + 1]
    return (float(lo), float(hi)) if as_float else (lo, hi)


def cpml_profiles_r(pml_width: List[int], fd_pad: List[int], dt: float,
                    grid_spacing: List[float], max_vel: float, pml_freq: float,
                    shape, dtype=np.float32) -> List[np.ndarray]:
    """``[a, b, db]`` per axis in shape order (z[, y], x), each 1-D of the
    runtime length of its axis."""
    out: List[np.ndarray] = []
    for dim in range(len(shape)):
        max_pml = max(pml_width[2 * dim] * grid_spacing[dim],
                      pml_width[2 * dim + 1] * grid_spacing[dim])
        a, b = ramp_profiles(pml_width[2 * dim:2 * dim + 2],
                             _band_starts(pml_width, fd_pad, shape, dim, True),
                             max_pml, dt, shape[dim], max_vel, dtype, pml_freq)
        db = np.gradient(b, grid_spacing[dim])
        out.extend([a, b, db])
    return out


def cpml_profiles_s(pml_width: List[int], fd_pad: List[int], dt: float,
                    grid_spacing: List[float], max_vel: float, pml_freq: float,
                    shape, dtype=np.float32) -> List[np.ndarray]:
    """``[a, b, ah, bh]`` per axis in shape order, ``ah, bh`` on the half cells."""
    out: List[np.ndarray] = []
    for dim in range(len(shape)):
        max_pml = max(pml_width[2 * dim] * grid_spacing[dim],
                      pml_width[2 * dim + 1] * grid_spacing[dim])
        starts = _band_starts(pml_width, fd_pad, shape, dim, False)
        w = pml_width[2 * dim:2 * dim + 2]
        a, b = ramp_profiles(w, starts, max_pml, dt, shape[dim], max_vel, dtype, pml_freq, start=0.0)
        ah, bh = ramp_profiles(w, starts, max_pml, dt, shape[dim], max_vel, dtype, pml_freq, start=0.5)
        out.extend([a, b, ah, bh])
    return out
