"""ORACLE -- test infrastructure only.

numpy restatement of the record wire format and of the L2 misfit, the
checker of ``swp_record_to_canonical`` / ``swp_record_to_raw`` /
``swp_l2_misfit``.  Only ``tests/`` import it.

* raw -> canonical: ``(B, nrec, nt)`` -> ``(B, nt, nrec, 1)`` and
  ``(nfield, B, nrec, nt)`` -> ``(B, nt, nrec, nfield)``
  (``propagator/_c.py:109-122``);
* canonical -> raw, the inverse (``_c.py:158-186``);
* the misfit 0.5 * sum(((pred - obs) * mask)^2): the residual in float32
  (the products the reference's tests form with torch), the sum in float64
  (``math.fsum`` of the float64 squares: exactly rounded, so any correctly
  accumulated float64 sum lies within n ulp of it).
"""
from __future__ import annotations

import math

import numpy as np


def to_canonical(raw: np.ndarray) -> np.ndarray:
    """_c.py:109-122: permute (0, 2, 1) + trailing field axis, or (1, 3, 2, 0)."""
    if raw.ndim == 3:
        return np.ascontiguousarray(raw.transpose(0, 2, 1)[..., None])
    if raw.ndim == 4:
        return np.ascontiguousarray(raw.transpose(1, 3, 2, 0))
    raise ValueError(raw.shape)


def to_raw(canonical: np.ndarray, raw_ndim: int) -> np.ndarray:
    """_c.py:158-186: the inverse permutation."""
    if raw_ndim == 3:
        g = canonical[..., 0] if canonical.ndim == 4 else canonical
        return np.ascontiguousarray(g.transpose(0, 2, 1))
    return np.ascontiguousarray(canonical.transpose(3, 0, 2, 1))


def residual(pred: np.ndarray, obs: np.ndarray, mask: np.ndarray | None) -> np.ndarray:
    """float32 (pred - obs) * mask[b, None, r, None]."""
    r = (pred.astype(np.float32) - obs.astype(np.float32)).astype(np.float32)
    if mask is not None:
        B, _, nrec, _ = pred.shape
        r = (r * mask.astype(np.float32).reshape(B, 1, nrec, 1)).astype(np.float32)
    return r


def l2_misfit(pred: np.ndarray, obs: np.ndarray, mask: np.ndarray | None = None) -> float:
    r = residual(pred, obs, mask).astype(np.float64).ravel()
    return 0.5 * math.fsum((r * r).tolist())
