"""Grid geometry and call-signature checks of the propagator (host side).

* :func:`normalize_io` restates ``core/validation.py:35-126`` (acquisition
  modes A1 / A2 / B, decided by the source rank; same error behaviour);
* :func:`runtime_coord_offset` restates ``core/geometry.py:48-56``: a physical
  (x[, y], z) coordinate moves by its axis' low-side pad plus the stencil
  halo ``M`` onto the runtime grid;
* :class:`EdgePadding` restates ``utils/torch.py:8-41``: replicate padding of a
  model whose backward crops (the padded cells' gradient is dropped, not
  folded into the edge cells);
* :func:`ricker` restates ``signal.py:10-21``.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F


def shape_tuple(value):
    shape = getattr(value, "shape", None)
    if shape is None:
        shape = np.shape(value)
    return tuple(int(d) for d in shape)


def normalize_io(wavelet, sources, receivers, ndim):
    """``(mode, batch, nsrc_per_shot, nrec, encoded)`` for the three modes."""
    ws, ss, rs = shape_tuple(wavelet), shape_tuple(sources), shape_tuple(receivers)
    if len(rs) != 3 or rs[-1] != ndim:
        raise ValueError(f"receivers must have shape (B, nrec, {ndim}); got {rs}. "
                         "Pre-broadcast/repeat per-shot if you previously passed a "
                         "shared (nrec, dim) array.")
    nrec = rs[1]
    if len(ss) == 2:
        if ss[-1] != ndim:
            raise ValueError(f"sources must have shape (nshots, {ndim}); got {ss}.")
        nshots = ss[0]
        if rs[0] != nshots:
            raise ValueError(f"receivers batch ({rs[0]}) must match sources nshots ({nshots}) "
                             "in naive multi-shot mode.")
        if len(ws) == 1:
            return "A1", nshots, 1, nrec, False
        if len(ws) == 2:
            if ws[0] != nshots:
                raise ValueError(f"wavelet must have shape (nshots={nshots}, nt); got {ws}.")
            return "A2", nshots, 1, nrec, False
        raise ValueError("wavelet must have shape (nt,) [shared] or (nshots, nt) [per-shot] in "
                         f"naive multi-shot mode; got {ws}.")
    if len(ss) == 3:
        if ss[0] != 1 or ss[-1] != ndim:
            raise ValueError(f"sources in source-encoding mode must have shape (1, nsrc, {ndim}); got {ss}.")
        nsrc = ss[1]
        if rs[0] != 1:
            raise ValueError(f"receivers batch must be 1 in source-encoding mode; got {rs[0]}.")
        if len(ws) == 1:
            return "B", 1, nsrc, nrec, True
        if len(ws) == 2:
            if ws[0] != nsrc:
                raise ValueError(f"wavelet must have shape (nt,) or (nsrc={nsrc}, nt) in "
                                 f"source-encoding mode; got {ws}.")
            return "B", 1, nsrc, nrec, True
        raise ValueError(f"wavelet must have shape (nt,) or (nsrc, nt) in source-encoding mode; got {ws}.")
    raise ValueError(f"sources must have shape (nshots, {ndim}) [naive multi-shot] or "
                     f"(1, nsrc, {ndim}) [source encoding]; got {ss}.")


def runtime_coord_offset(pad, halo, ndim):
    """Per-axis shift in (x[, y], z) order: low-side pad + halo."""
    return tuple(pad[2 * ax] + halo for ax in reversed(range(ndim)))


def runtime_shape(shape, pad, halo):
    """Physical shape -> runtime grid (pad on both faces + halo on both faces)."""
    return tuple(shape[ax] + pad[2 * ax] + pad[2 * ax + 1] + 2 * halo for ax in range(len(shape)))


def torch_pad_order(pad, ndim, halo):
    """``F.pad`` pairs (last axis first) of pad + halo per face."""
    out = []
    for ax in reversed(range(ndim)):
        out += [pad[2 * ax] + halo, pad[2 * ax + 1] + halo]
    return out


class EdgePadding(torch.autograd.Function):
    """Replicate padding whose backward crops the padded cells away."""

    @staticmethod
    def forward(ctx, x, pad):
        ctx.pad = pad
        return F.pad(x.unsqueeze(0), pad=pad, mode="replicate").squeeze(0)

    @staticmethod
    def backward(ctx, g):
        sl = [slice(None)] * g.dim()
        for i in range(len(ctx.pad) // 2):
            lo, hi = ctx.pad[2 * i], ctx.pad[2 * i + 1]
            sl[-(i + 1)] = slice(lo, -hi if hi > 0 else None)
        return g[tuple(sl)], None


def crop_physical(a, pad, halo, ndim):
    """Runtime-grid array -> physical box (the last ``ndim`` axes)."""
    sl = [Ellipsis]
    for ax in range(ndim):
        lo, hi = pad[2 * ax] + halo, pad[2 * ax + 1] + halo
        sl.append(slice(lo, -hi if hi > 0 else None))
    return a[tuple(sl)]


def ricker(t, f=10.0):
    """Ricker wavelet ``(1 - 2 (pi f t)^2) exp(-(pi f t)^2)`` as float32."""
    r = (1 - 2 * (np.pi * f * t) ** 2) * np.exp(-(np.pi * f * t) ** 2)
    return r.astype(np.float32)
