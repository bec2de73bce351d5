"""Autotuner for the launch shapes of a gradient workload: shots per launch
and the z-marching kernels' segment heights.

Each candidate is timed on a short probe (the first ``probe_nt`` steps of
the workload, forward + backward of one launch, wall clock around a
synchronised run, best of ``repeats`` after one warm-up); the per-step cost
does not depend on nt.  The search is coordinate-wise: shots per launch
first with the default segment heights, then each knob in turn with the
others at their best.  A shots-per-launch candidate is skipped when its
full-length workspace and tape (``swp_query`` at the workload's nt) would
not fit in ``max_bytes``.  The chosen heights are installed in the core's
tuning table (``swp_set_tuning``).  3-D records and gradients do not depend on
them.  2-D gradients move at the float32 rounding level (4e-8 relative
measured) with the reconstruction segment height, and 2-D records do too
(1.2e-6 measured at 96 x 128) where a change of the interior segment height
moves cells between the interior and band kernels.  The 2-D heights act only
on launches that run the z-marching kernels (from 4M shot-cells up).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from . import _core
from .equations import Acoustic, Acoustic3D
from .loss import l2_misfit

KNOBS = {
    "acoustic2d": {_core.TUNE_TZ2: (16, 32, 64), _core.TUNE_TZ2_BAND: (4, 8), _core.TUNE_TZ2_RECON: (16, 32)},
    "acoustic3d": {_core.TUNE_TZ3: (32, 48, 64)},
    "elastic2d": {},
}
KNOB_NAMES = {_core.TUNE_TZ2: "tz2", _core.TUNE_TZ2_BAND: "tz2_band", _core.TUNE_TZ2_RECON: "tz2_recon",
              _core.TUNE_TZ3: "tz3"}


@dataclass
class TuneResult:
    shots_per_launch: int
    knobs: dict                                  # name -> value installed
    seconds_per_shot_step: float                 # best probe time per shot and time step
    table: list = field(default_factory=list)    # (shots_per_launch, knobs, seconds per shot-step)


def _family(solver):
    eq = solver.equation
    if isinstance(eq, Acoustic3D):
        return "acoustic3d"
    if isinstance(eq, Acoustic):
        return "acoustic2d"
    return "elastic2d"


def _probe(solver, wav, src, rec, models, per_launch, repeats):
    """Seconds per shot and time step of one launch of per_launch shots."""
    ms = [m.detach().clone().requires_grad_(True) for m in models]
    best = float("inf")
    nt = int(wav.shape[-1])
    for k in range(repeats + 1):
        for m in ms:
            m.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pred = solver(wav, src[:per_launch], rec[:per_launch], models=ms)
        l2_misfit(pred, torch.zeros_like(pred)).backward()
        torch.cuda.synchronize()
        if k > 0:
            best = min(best, (time.perf_counter() - t0) / (per_launch * nt))
    return best


def _fits(solver, per_launch, full_nt, max_bytes):
    call = getattr(solver, "_last_call", None)
    if call is None:
        return True
    P = _core.SwpProblem.from_buffer_copy(call.problem)
    P.B = per_launch
    P.nt = full_nt
    names = call.names
    ws_f, tape = _core.query(names[0], P)
    entry = {_core.MEM_FULL: 1, _core.MEM_BOUNDARY: 2, _core.MEM_CKPT: 3, _core.MEM_BINOMIAL: 4}.get(call.memory, 2)
    ws_b, _ = _core.query(names[entry], P)
    return max(ws_f, ws_b) + tape <= max_bytes


def tune(solver, wavelet, sources, receivers, models, nshots: int, full_nt: int | None = None,
         per_launch_candidates=None, probe_nt: int = 60, repeats: int = 2, max_bytes: int | None = None,
         install: bool = True) -> TuneResult:
    """Time the candidates and return (and, with install, apply) the best.

    wavelet: (nt,) or what the solver takes for this batch layout, truncated
    here to probe_nt steps; sources / receivers: per-shot arrays (at least
    the largest candidate's shots); models: the model tensors."""
    fam = _family(solver)
    wav = torch.as_tensor(wavelet)[..., :probe_nt].contiguous()
    full_nt = int(full_nt or torch.as_tensor(wavelet).shape[-1])
    if max_bytes is None:
        free, _ = torch.cuda.mem_get_info()
        max_bytes = int(0.85 * free)
    if per_launch_candidates is None:
        per_launch_candidates = sorted({c for c in (1, 2, 4, 8, 16, 32, 64) if c <= nshots} | {nshots})
    knobs = KNOBS[fam]
    for key in knobs:
        _core.set_tuning(key, 0)
    table = []
    best_pl, best_t = None, float("inf")
    for pl in per_launch_candidates:
        if best_pl is not None and not _fits(solver, pl, full_nt, max_bytes):
            continue
        t = _probe(solver, wav, sources, receivers, models, pl, repeats)
        table.append((pl, {}, t))
        if not _fits(solver, pl, full_nt, max_bytes):
            continue
        if t < best_t:
            best_pl, best_t = pl, t
    if best_pl is None:
        raise RuntimeError(f"no shots-per-launch candidate of {list(per_launch_candidates)} fits in "
                           f"{max_bytes} bytes at nt = {full_nt}")
    chosen = {}
    for key, values in knobs.items():
        best_v = None
        for v in values:
            _core.set_tuning(key, v)
            t = _probe(solver, wav, sources, receivers, models, best_pl, repeats)
            table.append((best_pl, {**{KNOB_NAMES[k]: x for k, x in chosen.items()}, KNOB_NAMES[key]: v}, t))
            if t < best_t or best_v is None and t <= best_t:
                best_t, best_v = t, v
        chosen[key] = best_v if best_v is not None else 0
        _core.set_tuning(key, chosen[key])
    if not install:
        for key in knobs:
            _core.set_tuning(key, 0)
    return TuneResult(best_pl, {KNOB_NAMES[k]: v for k, v in chosen.items()}, best_t, table)
