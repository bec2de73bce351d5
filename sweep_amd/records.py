"""Receiver records between the core's raw layout and the canonical one.

The forward entries write records raw, time innermost: ``(B, nrec, nt)`` for
one field, ``(nfield, B, nrec, nt)`` for several.  Callers, losses and the
reference's other backends use the canonical ``(B, nt, nrec, nfield)``
(``propagator/_c.py:95-186``).  Both directions run as one HIP transpose
(``swp_record_to_canonical`` / ``swp_record_to_raw``); there is no host path.
"""
from __future__ import annotations

import torch

from . import _core


def _check(t: torch.Tensor, what: str):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError(f"{what} must be a float32 tensor on the GPU")


def to_canonical(raw: torch.Tensor) -> torch.Tensor:
    """``(B, nrec, nt)`` -> ``(B, nt, nrec, 1)``; ``(nfield, B, nrec, nt)`` ->
    ``(B, nt, nrec, nfield)`` (``_c.py:109-122``)."""
    _check(raw, "record")
    if raw.ndim == 3:
        nf, (B, nrec, nt) = 1, raw.shape
    elif raw.ndim == 4:
        nf, B, nrec, nt = raw.shape
    else:
        raise ValueError(f"Unexpected raw record ndim={raw.ndim}, shape={tuple(raw.shape)}; "
                         "expected (B, nrec, nt) or (nfield, B, nrec, nt).")
    raw = raw.contiguous()
    out = torch.empty((B, nt, nrec, nf), dtype=torch.float32, device=raw.device)
    L = _core.load()
    _core.on_stream(raw.device, L.swp_record_to_canonical, raw.data_ptr(), out.data_ptr(), nf, B, nrec, nt)
    return out


def to_raw(canonical: torch.Tensor, raw_ndim: int) -> torch.Tensor:
    """Inverse of :func:`to_canonical` (``_c.py:158-186``): ``raw_ndim`` 3
    gives ``(B, nrec, nt)`` (a 3-D ``(B, nt, nrec)`` input is accepted), 4
    gives ``(nfield, B, nrec, nt)``."""
    _check(canonical, "record")
    g = canonical
    if raw_ndim == 3:
        if g.ndim == 4:
            if g.shape[-1] != 1:
                raise ValueError("Single-channel backward expects canonical grad with last dim = 1, "
                                 f"got shape {tuple(g.shape)}.")
            g = g[..., 0]
        if g.ndim != 3:
            raise ValueError(f"Single-channel backward needs 3-D grad, got {tuple(g.shape)}.")
        B, nt, nrec = g.shape
        nf = 1
        shape = (B, nrec, nt)
    elif raw_ndim == 4:
        if g.ndim != 4:
            raise ValueError(f"Multi-channel backward needs 4-D grad, got {tuple(g.shape)}.")
        B, nt, nrec, nf = g.shape
        shape = (nf, B, nrec, nt)
    else:
        raise ValueError(f"Unexpected raw ndim={raw_ndim}; expected 3 or 4.")
    g = g.contiguous()
    out = torch.empty(shape, dtype=torch.float32, device=g.device)
    L = _core.load()
    _core.on_stream(g.device, L.swp_record_to_raw, g.data_ptr(), out.data_ptr(), nf, B, nrec, nt)
    return out
