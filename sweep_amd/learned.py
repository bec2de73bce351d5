"""Learned-preconditioner FWI: a U-Net (``sweep_amd.unet``) maps
[gradient, source illumination, current vp] to a model update, trained on
pairs the solver itself generates; the plain illumination-preconditioned
line-search step (``sweep_amd.fwi.PlainFWI``) is both the fallback and the
arbiter (a network step is kept only if the solver's misfit falls).

* Pairs: for seeded random layered models, plain FWI runs from a smoothed
  start; at each iteration the input planes (g, S, vp) and the target
  v_true - vp are recorded.
* Scaling, per sample: g / max|g|, S / max S, (vp - 3000) / 1500 in; the
  target in units of ``TARGET_SCALE`` m/s out.
* A learned iteration computes g and S as the plain one does, evaluates the
  network, and line-searches the step along its update (scales 1, 1/2,
  1/4); the step is taken only if the solver's misfit falls, otherwise the
  iteration falls back to the plain line search along the preconditioned
  gradient.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from .fwi import IterationRecord, PlainFWI, line_search
from .unet import UNet

TARGET_SCALE = 100.0   # m/s per unit of network output
VP_CENTER, VP_SCALE = 3000.0, 1500.0


def network_planes(g: torch.Tensor, S: torch.Tensor, vp: torch.Tensor):
    """The three scaled (H, W) input planes and their scales (1 each)."""
    gs = g / g.abs().max().clamp_min(1e-30)
    ss = S / S.max().clamp_min(1e-30)
    vs = (vp - VP_CENTER) / VP_SCALE
    return [gs.float(), ss.float(), vs.float()]


@dataclass
class Pair:
    planes: list          # 3 float (H, W) tensors on the GPU
    target: torch.Tensor  # (H, W) float, (v_true - vp) / TARGET_SCALE


def generate_pairs(make_fwi, models, iterations: int, device, log=None) -> list:
    """make_fwi(vp_true) -> (PlainFWI, vp_start); records `iterations` pairs
    along each model's plain FWI path (log(k, n) after each model)."""
    pairs = []
    for k, vt in enumerate(models):
        if log is not None and k > 0:
            log(k, len(models))
        fwi, vp = make_fwi(vt)
        vt = torch.as_tensor(vt, device=device)
        for _ in range(iterations):
            start = vp
            vp = fwi.iterate(vp)            # computes g and S at `start` on the way
            _, g, S = fwi.last
            pairs.append(Pair(network_planes(g, S, start), ((vt - start) / TARGET_SCALE).float().contiguous()))
    return pairs


def train(net: UNet, pairs, steps: int, batch: int = 2, lr: float = 1e-3, seed: int = 0):
    """Adam on mean-squared error over random mini-batches; returns the losses."""
    rng = np.random.default_rng(seed)
    losses = []
    for _ in range(steps):
        idx = rng.choice(len(pairs), size=min(batch, len(pairs)), replace=False)
        planes = [torch.stack([pairs[i].planes[c] for i in idx]) for c in range(3)]
        target = torch.stack([pairs[i].target for i in idx])
        x, _ = net.input(planes, [1.0, 1.0, 1.0])
        losses.append(net.train_step(x, target, lr=lr))
    return [float(l) for l in losses]


class LearnedFWI(PlainFWI):
    """PlainFWI whose iterations try the network's update first."""

    def __init__(self, net: UNet, *args, learned_scales=(1.0, 0.5, 0.25), **kw):
        super().__init__(*args, **kw)
        self.net = net
        self.learned_scales = tuple(learned_scales)
        self.learned_accepted = []

    def update(self, g, S, vp) -> torch.Tensor:
        x, _ = self.net.input([p[None] for p in network_planes(g, S, vp)], [1.0, 1.0, 1.0])
        out = self.net.forward(x)
        return out[0, ..., 0] * TARGET_SCALE

    def iterate(self, vp: torch.Tensor) -> torch.Tensor:
        torch.cuda.synchronize(vp.device)
        t0 = time.perf_counter()
        J0, g, S = self.gradient(vp)
        self.last = (J0, g, S)
        du = self.update(g, S, vp)
        evals = 0
        best = None
        for s in self.learned_scales:
            cand = (vp + s * du).clamp_(self.vmin, self.vmax)
            J = self.misfit(cand)
            evals += 1
            if J < J0:
                best = (cand, J, s)
                break
        if best is not None:
            out, J1, step = best
            self.learned_accepted.append(True)
        else:
            self.learned_accepted.append(False)
            d = self.direction(g, S)
            res = line_search(lambda st: self.misfit(self.trial(vp, d, st)), J0, self.next_step,
                              self.max_halvings, self.max_doublings)
            evals += res.evals
            out, J1, step = vp, J0, 0.0
            if res.step > 0:
                out = self.trial(vp, d, res.step)
                J1, step = res.misfit, res.step
                self.next_step = res.step
        torch.cuda.synchronize(vp.device)
        self.history.append(IterationRecord(J0, J1, step, evals, time.perf_counter() - t0))
        return out


def _main():
    """Train on solver-generated pairs at a grid (default 384 x 1152, the
    2-D workload's), then run plain and learned FWI from the same start on a
    held-out model; one JSON line per stage and per iteration."""
    import argparse
    import json

    from . import synthetic
    from .equations import Acoustic
    from .options import BoundarySaving
    from .propagator import PropTorch

    ap = argparse.ArgumentParser()
    ap.add_argument("--nz", type=int, default=384)
    ap.add_argument("--nx", type=int, default=1152)
    ap.add_argument("--nt", type=int, default=3000)
    ap.add_argument("--shots", type=int, default=32)
    ap.add_argument("--pair-shots", type=int, default=16,
                    help="shots per gradient when generating training pairs")
    ap.add_argument("--train-models", type=int, default=64)
    ap.add_argument("--pair-iterations", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--iterations", type=int, default=8)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--learned-scales", default="1,0.5,0.25",
                    help="comma-separated trial scales of the network's update, tried in order")
    ap.add_argument("--stage", choices=["all", "train", "compare"], default="all",
                    help="all: pairs, training and the comparison; train: pairs from --first-model on and "
                         "training (from --load if given), weights to --save; compare: --load, then plain and "
                         "learned FWI")
    ap.add_argument("--first-model", type=int, default=0, help="index of the first training model (seed 100 + k)")
    ap.add_argument("--load", default=None, help="network weights and Adam state to start from")
    ap.add_argument("--save", default=None, help="where to write the trained network")
    args = ap.parse_args()
    scales = tuple(float(v) for v in args.learned_scales.split(","))
    dev = torch.device("cuda:0")
    shape, dh, dt, f0 = (args.nz, args.nx), 8.0, 8e-4, 8.0
    wav = synthetic.ricker(np.arange(args.nt) * dt - 1.5 / f0, f0)
    solver = PropTorch(Acoustic(spatial_order=8, device=dev), shape=shape, dh=dh, dt=dt, device=dev, abcn=50,
                       memory=BoundarySaving())
    def acquisition(nshots):
        src, rec, mask = synthetic.surface_acquisition(shape, nshots, seed=0)
        batches = [(a, min(a + 32, nshots)) for a in range(0, nshots, 32)]
        return src, rec, batches, [torch.tensor(mask[a:b], device=dev) for a, b in batches]

    acq = {n: acquisition(n) for n in {args.shots, args.pair_shots or args.shots}}

    def make(vt_np, cls=PlainFWI, extra=(), nshots=args.shots, **kw):
        src, rec, batches, masks = acq[nshots]
        vt = torch.tensor(vt_np, device=dev)
        with torch.no_grad():
            obs = [solver(wav, src[a:b], rec[a:b], models=[vt]) for a, b in batches]
        f = cls(*extra, solver, wav, src, rec, obs, batches, masks, vmin=1400.0, vmax=4600.0, first_step=50.0,
                **kw)
        return f, torch.tensor(synthetic.smoothed(vt_np), device=dev)

    net = UNet(dev, levels=4, base=args.base, nsrc=3, seed=0)
    if args.load:
        net.load(args.load)
    if args.stage in ("all", "train"):
        t0 = time.perf_counter()
        train_models = [synthetic.layered_2d(shape, seed=100 + k)
                        for k in range(args.first_model, args.first_model + args.train_models)]
        pairs = generate_pairs(lambda vt: make(vt, nshots=args.pair_shots or args.shots), train_models,
                               args.pair_iterations, dev,
                               log=lambda k, n: print(json.dumps({"stage": "pairs", "models_done": k, "of": n,
                                                                  "seconds": round(time.perf_counter() - t0, 1)}),
                                                      flush=True))
        torch.cuda.synchronize()
        print(json.dumps({"stage": "pairs", "count": len(pairs), "seconds": round(time.perf_counter() - t0, 2)}),
              flush=True)
        t0 = time.perf_counter()
        losses = train(net, pairs, args.train_steps, batch=args.batch, lr=args.lr, seed=args.first_model)
        torch.cuda.synchronize()
        print(json.dumps({"stage": "train", "steps": args.train_steps, "loss_first": losses[0],
                          "loss_last": float(np.mean(losses[-20:])), "seconds": round(time.perf_counter() - t0, 2)}),
              flush=True)
        del pairs
        if args.save:
            net.save(args.save)
        if args.stage == "train":
            return
    held = synthetic.layered_2d(shape, seed=7)
    for name, cls, extra, kw in (("plain", PlainFWI, (), {}),
                                 ("learned", LearnedFWI, (net,), {"learned_scales": scales})):
        f, vp = make(held, cls, extra, **kw)
        vt = torch.tensor(held, device=dev)
        wall = 0.0
        for k in range(args.iterations):
            vp = f.iterate(vp)
            r = f.history[-1]
            wall += r.seconds
            rec_ = {"run": name, "iteration": k, "misfit_after": r.misfit_after, "step": r.step, "evals": r.evals,
                    "wall": round(wall, 3), "model_rel_err": float((vp - vt).norm() / vt.norm())}
            if name == "learned":
                rec_["learned_step"] = f.learned_accepted[-1]
            print(json.dumps(rec_), flush=True)


if __name__ == "__main__":
    _main()
