"""Plain full-waveform inversion on the propagator: the illumination-
preconditioned gradient step with a line search.

One iteration over the shot batches:

1. the misfit J(m) = sum over batches of 0.5 ||mask (pred - obs)||^2
   (float64, ``sweep_amd.loss.l2_misfit``) and its gradient g, with the
   source illumination S = sum_t u_tt^2 summed over shots (the reference's
   ``compute_illumination``, ``propagator/_c.py:1040-1064``: the pseudo-Hessian
   paired with the gradient, so g / (S + eps) is a preconditioned gradient);
2. the direction d = -(g / (S + eps * max S)), scaled so max |d| = 1 m/s;
3. a line search along d over the step (m/s), models clipped to
   [vmin, vmax]: the first trial is the last accepted step; a trial that
   lowers J is doubled while that keeps lowering it, one that does not is
   halved (at most ``max_halvings`` times).  A step is taken only if the
   solver's misfit falls; if none does, the iteration reports no progress
   and the model stays.

The reference ships no FWI driver (its examples build one in notebooks);
this one is the plain method against which a learned step is measured: a
learned update is accepted only where it lowers the same misfit.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Sequence

import torch

from .loss import l2_misfit


@dataclass
class LineSearchResult:
    step: float          # accepted step (0: none)
    misfit: float        # J at the accepted step (J0 when none)
    evals: int           # misfit evaluations spent
    trials: list = field(default_factory=list)   # (step, J) in order


def line_search(f: Callable[[float], float], f0: float, step0: float, max_halvings: int = 5,
                max_doublings: int = 3) -> LineSearchResult:
    """Backtracking / expanding search for a step that lowers f below f0.

    ``f(step)`` is the misfit at the trial model.  Host logic only (no GPU)."""
    if not step0 > 0:
        raise ValueError("the first trial step must be positive")
    trials = []
    s = float(step0)
    fs = f(s)
    trials.append((s, fs))
    if fs < f0:
        for _ in range(max_doublings):
            s2 = 2.0 * s
            f2 = f(s2)
            trials.append((s2, f2))
            if not f2 < fs:
                break
            s, fs = s2, f2
        return LineSearchResult(s, fs, len(trials), trials)
    for _ in range(max_halvings):
        s *= 0.5
        fs = f(s)
        trials.append((s, fs))
        if fs < f0:
            return LineSearchResult(s, fs, len(trials), trials)
    return LineSearchResult(0.0, f0, len(trials), trials)


@dataclass
class IterationRecord:
    misfit: float        # J before the step
    misfit_after: float  # J after the step (== misfit when rejected)
    step: float          # m/s along the unit direction (0: rejected)
    evals: int           # forward-only misfit evaluations in the line search
    seconds: float       # wall time of the iteration (GPU synchronised)


class PlainFWI:
    """Illumination-preconditioned line-search FWI over shot batches.

    solver: a ``PropTorch`` acoustic propagator; wavelet: (nt,) or what the
    solver accepts; sources / receivers: per-shot arrays indexed by batch
    slices; observed: one canonical record tensor per batch; mask: per-batch
    (B, nrec) weights or None; batches: [(a, b), ...] shot slices."""

    def __init__(self, solver, wavelet, sources, receivers, observed: Sequence[torch.Tensor],
                 batches: Sequence[tuple], masks: Sequence | None = None, vmin: float = 1400.0,
                 vmax: float = 5000.0, eps: float = 1e-3, first_step: float = 50.0,
                 max_halvings: int = 5, max_doublings: int = 2):
        self.solver = solver
        self.wavelet = wavelet
        self.sources = sources
        self.receivers = receivers
        self.observed = list(observed)
        self.batches = list(batches)
        self.masks = list(masks) if masks is not None else [None] * len(self.batches)
        if not (len(self.observed) == len(self.batches) == len(self.masks)):
            raise ValueError("observed, batches and masks must have one entry per batch")
        self.vmin, self.vmax = float(vmin), float(vmax)
        self.eps = float(eps)
        self.next_step = float(first_step)
        self.max_halvings = max_halvings
        self.max_doublings = max_doublings
        self.history: list[IterationRecord] = []
        self.last = None   # (J, g, S) of the last iteration's starting model

    # -- objective -------------------------------------------------------
    def misfit(self, vp: torch.Tensor) -> float:
        """J(vp), forward only."""
        total = torch.zeros((), dtype=torch.float64, device=vp.device)
        with torch.no_grad():
            for (a, b), obs, m in zip(self.batches, self.observed, self.masks):
                pred = self.solver(self.wavelet, self.sources[a:b], self.receivers[a:b], models=[vp])
                total += l2_misfit(pred, obs, m)
        return float(total)

    def gradient(self, vp: torch.Tensor):
        """(J, g, S): misfit, model gradient and source illumination at vp."""
        v = vp.detach().clone().requires_grad_(True)
        total = torch.zeros((), dtype=torch.float64, device=vp.device)
        illum = torch.zeros_like(vp)
        prev = getattr(self.solver, "compute_illumination", False)
        self.solver.compute_illumination = True
        try:
            for (a, b), obs, m in zip(self.batches, self.observed, self.masks):
                pred = self.solver(self.wavelet, self.sources[a:b], self.receivers[a:b], models=[v])
                loss = l2_misfit(pred, obs, m)
                loss.backward()
                total += loss.detach()
                S, _ = self.solver.collect_illumination()
                illum += S
        finally:
            self.solver.compute_illumination = prev
        return float(total), v.grad.detach(), illum

    def direction(self, g: torch.Tensor, illum: torch.Tensor) -> torch.Tensor:
        """-(g / (S + eps max S)), scaled to max |d| = 1."""
        d = -g / (illum + self.eps * illum.max().clamp_min(1e-30))
        peak = d.abs().max()
        if not bool(peak > 0):
            return torch.zeros_like(d)
        return d / peak

    def trial(self, vp: torch.Tensor, d: torch.Tensor, step: float) -> torch.Tensor:
        return (vp + step * d).clamp_(self.vmin, self.vmax)

    # -- iteration --------------------------------------------------------
    def iterate(self, vp: torch.Tensor) -> torch.Tensor:
        """One FWI iteration; returns the new model (vp itself when no step lowers J)."""
        torch.cuda.synchronize(vp.device)
        t0 = time.perf_counter()
        J0, g, S = self.gradient(vp)
        self.last = (J0, g, S)
        d = self.direction(g, S)
        res = line_search(lambda s: self.misfit(self.trial(vp, d, s)), J0, self.next_step,
                          self.max_halvings, self.max_doublings)
        out = vp
        if res.step > 0:
            out = self.trial(vp, d, res.step)
            self.next_step = res.step
        else:
            self.next_step = max(self.next_step * 0.25, 1e-3)
        torch.cuda.synchronize(vp.device)
        self.history.append(IterationRecord(J0, res.misfit, res.step, res.evals, time.perf_counter() - t0))
        return out

    def run(self, vp: torch.Tensor, iterations: int) -> torch.Tensor:
        for _ in range(iterations):
            vp = self.iterate(vp)
        return vp


def _main():
    """Plain FWI at the 2-D workload's grid (384 x 1152, dh 8 m, dt 0.8 ms,
    order 8, seeded sinuous layers, 64 surface shots of 150-500 receivers)
    from a smoothed start; one JSON line per iteration."""
    import argparse
    import json

    import numpy as np

    from . import synthetic
    from .equations import Acoustic
    from .options import BoundarySaving
    from .propagator import PropTorch

    ap = argparse.ArgumentParser()
    ap.add_argument("--iterations", type=int, default=5)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--per-launch", type=int, default=32)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    shape, dh, dt, f0 = (384, 1152), 8.0, 8e-4, 8.0
    vp_true = synthetic.layered_2d(shape, seed=args.seed)
    vp0 = synthetic.smoothed(vp_true)
    src, rec, mask = synthetic.surface_acquisition(shape, args.shots, seed=args.seed)
    wav = synthetic.ricker(np.arange(args.nt) * dt - 1.5 / f0, f0)
    solver = PropTorch(Acoustic(spatial_order=8, device=dev), shape=shape, dh=dh, dt=dt, device=dev, abcn=50,
                       memory=BoundarySaving())
    batches = [(a, min(a + args.per_launch, args.shots)) for a in range(0, args.shots, args.per_launch)]
    vt = torch.tensor(vp_true, device=dev)
    with torch.no_grad():
        obs = [solver(wav, src[a:b], rec[a:b], models=[vt]) for a, b in batches]
    masks = [torch.tensor(mask[a:b], device=dev) for a, b in batches]
    fwi = PlainFWI(solver, wav, src, rec, obs, batches, masks, vmin=1400.0, vmax=4600.0, first_step=50.0)
    vp = torch.tensor(vp0, device=dev)
    err0 = float((vp - vt).norm() / vt.norm())
    print(json.dumps({"grid": list(shape), "shots": args.shots, "nt": args.nt, "model_rel_err": err0}), flush=True)
    for k in range(args.iterations):
        vp = fwi.iterate(vp)
        r = fwi.history[-1]
        print(json.dumps({"iteration": k, "misfit": r.misfit, "misfit_after": r.misfit_after, "step": r.step,
                          "evals": r.evals, "seconds": round(r.seconds, 3),
                          "model_rel_err": float((vp - vt).norm() / vt.norm())}), flush=True)


if __name__ == "__main__":
    _main()
