"""Benchmark: 2-D acoustic FWI gradients (forward + adjoint, boundary saving).

Workload (one "step" = one gradient over the whole shot batch):
  grid 384 x 1152 at dh = 8 m, dt = 0.8 ms, nt = 3000, 8th order in space,
  CPML 50 cells, seeded sinuous-layer vp (1500-4500 m/s, +-5% smoothed noise),
  32 surface shots (Ricker 8 Hz, delay 1.5/f) with uneven receiver spreads
  of 150-500 receivers padded to the largest count with a residual mask;
  loss 0.5 ||mask (pred - obs)||^2 against the record of the true model,
  gradient w.r.t. the smoothed starting model by boundary saving.

Prints ONE JSON line: shot-gradients/s (value = 32 x steps / time), the
roofline of the dominant kernel (algorithmic bytes per launch over the mean
launch duration, measured with HIP events on the launch stream during the
timed steps), and the oracle's CPU time for a bounded sample of the same
workload on this machine's host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(name="acoustic2d_384x1152_nt3000_32shots_bs", shape=(384, 1152), dh=8.0, dt=8e-4,
                nt=3000, order=8, abcn=50, nshots=32, f0=8.0, seed=2)


def make_problem(dev, nt=None, nshots=None):
    import torch
    from sweep_amd import synthetic
    from sweep_amd.equations import Acoustic
    from sweep_amd.options import BoundarySaving
    from sweep_amd.propagator import PropTorch
    W = WORKLOAD
    nt = nt or W["nt"]
    nshots = nshots or W["nshots"]
    vp_true = synthetic.layered_2d(W["shape"], seed=W["seed"])
    vp_init = synthetic.smoothed(vp_true)
    src, rec, mask = synthetic.surface_acquisition(W["shape"], nshots, seed=W["seed"])
    t = np.arange(nt) * W["dt"]
    wav = synthetic.ricker(t - 1.5 / W["f0"], W["f0"])
    solver = PropTorch(Acoustic(spatial_order=W["order"], device=dev), shape=W["shape"], dh=W["dh"],
                       dt=W["dt"], device=dev, abcn=W["abcn"], memory=BoundarySaving())
    mask_t = torch.tensor(mask, device=dev)[:, None, :, None]
    return solver, vp_true, vp_init, src, rec, mask_t, wav


def algorithmic_bytes(B, nz, nx, M, pad, W):
    """Bytes per launch of the three stencil kernels under the counting rule:
    4 B per distinct full-grid float read or written per cell, perfect halo
    reuse; +16 B per cell per axis inside that axis' CPML compute band (psi,
    zeta read, next psi, zeta written); +4 B per saved strip cell; +12 B per
    imaged cell (lambda read, gradient read + write)."""
    iz = np.arange(M, nz - M)[:, None]
    ix = np.arange(M, nx - M)[None, :]
    cells = iz.size * ix.size
    cx = ((ix < pad + 2 * M) | (ix >= nx - pad - 2 * M)).sum() * iz.size
    cz = ((iz < pad + 2 * M) | (iz >= nz - pad - 2 * M)).sum() * ix.size
    nxp, nzp = nx - 2 * pad - 2 * M, nz - 2 * pad - 2 * M
    strip = 2 * W * nxp + 2 * W * nzp
    fwd = B * (16 * cells + 16 * (cx + cz) + 4 * strip)
    adj = B * (16 * cells + 16 * (cx + cz))
    lo = pad + 2 * M + 1
    rcells = (nz - 2 * lo) * (nx - 2 * lo)
    img = (nzp - 2 * W) * (nxp - 2 * W)
    recon = B * (16 * rcells + 12 * img)
    return {0: fwd, 1: adj, 2: recon}


def cpu_baseline(sample_nt=1000):
    """The oracle (C, OpenMP) on one shot of the same grid for sample_nt steps
    of forward + boundary-saving backward, extrapolated to the full nt."""
    from oracle.oracle import A2DProblem
    from sweep_amd import synthetic
    W = WORKLOAD
    P = A2DProblem(W["order"], W["shape"], W["dh"], W["dt"], sample_nt, abcn=W["abcn"], B=1)
    vp = P.pad_model(synthetic.smoothed(synthetic.layered_2d(W["shape"], seed=W["seed"])))
    t = np.arange(sample_nt) * W["dt"]
    wav = np.ascontiguousarray(synthetic.ricker(t - 1.5 / W["f0"], W["f0"])[None, None, :])
    src, rec, _ = synthetic.surface_acquisition(W["shape"], 1, seed=W["seed"], nrec_range=(300, 300))
    sl, rl = P.shift(src[:, None, :]), P.shift(rec)
    res = np.random.default_rng(0).standard_normal((1, rl.shape[1], sample_nt)).astype(np.float32)
    t0 = time.perf_counter()
    P.gradient(1, vp, wav, sl, rl, res)
    dt = time.perf_counter() - t0
    per_grad = dt * W["nt"] / sample_nt
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / per_grad, "unit": "shot-gradients/s", "cores": cores, "kind": "port",
            "sample": f"1 shot, {sample_nt} of {W['nt']} steps of forward + boundary-saving backward "
                      f"on the full 492x1260 runtime grid ({dt:.1f} s), extrapolated x{W['nt'] // sample_nt}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--timing-stride", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--nt", type=int, default=None, help="override nt (profiling runs only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from sweep_amd import _core
    dev = torch.device("cuda:0")
    solver, vp_true, vp_init, src, rec, mask, wav = make_problem(dev, nt=args.nt)
    W = WORKLOAD
    nshots = src.shape[0]
    print(f"[bench] {W['name']}: building obs", file=sys.stderr, flush=True)
    with torch.no_grad():
        obs = solver(wav, src, rec, models=[torch.tensor(vp_true, device=dev)])
    vp0 = torch.tensor(vp_init, device=dev)

    def step():
        vp = vp0.clone().requires_grad_(True)
        pred = solver(wav, src, rec, models=[vp])
        loss = 0.5 * ((pred - obs) * mask).pow(2).sum()
        loss.backward()
        return vp.grad

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    _core.set_timing(args.timing_stride)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        g = step()
        print(f"[bench] step {k + 1}/{args.steps} launched", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stats = {c: _core.kernel_stats(c) for c in range(4)}
    _core.set_timing(0)
    assert torch.isfinite(g).all(), "non-finite gradient"

    P = solver._last_call.problem
    abytes = algorithmic_bytes(nshots, P.nz, P.nx, P.M, W["abcn"], P.M + 1)
    names = {0: "fwd_kernel (forward step + strip save)", 1: "adj_kernel (exact adjoint step)",
             2: "recon_kernel (BS reconstruction + fused imaging)"}
    per = {}
    for c in (0, 1, 2):
        ms, n = stats[c]
        if n:
            per[c] = dict(avg_ms=ms / n, total_est_ms=ms * args.timing_stride,
                          gbs=abytes[c] / (ms / n * 1e-3) / 1e9)
    dom = max(per, key=lambda c: per[c]["total_est_ms"])
    ach = per[dom]["gbs"]
    nt = P.nt
    value = nshots * args.steps / elapsed
    out = {
        "metric": "shot-gradients/s fwd+adj", "value": value, "unit": "shot-gradients/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": W["name"] if args.nt is None else f"{W['name']}_nt{nt}",
                   "grid": list(W["shape"]), "runtime_grid": [P.nz, P.nx], "nt": nt, "shots": nshots,
                   "order": W["order"], "memory": "boundary_saving"},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": int(abytes[dom]),
                     "avg_launch_ms": per[dom]["avg_ms"],
                     "kernels": {names[c]: {"avg_launch_ms": per[c]["avg_ms"], "GB/s": per[c]["gbs"]}
                                 for c in per}},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
