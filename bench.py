"""Benchmark: acoustic FWI gradients (forward + adjoint, boundary saving).

Default workload, ``acoustic3d`` (one "step" = one launch: the gradients of
``--shots-per-launch`` shots, default 2, taken in turn from the 8 shots):
  grid (nz, ny, nx) = 256 x 512 x 512 at dh = 10 m, dt = 1 ms, nt = 2000,
  8th order in space, CPML 50 cells (runtime grid 364 x 620 x 620), seeded
  dipping-layer vp (1500-4000 m/s), 8 shots each recorded by a 64 x 64
  surface receiver patch; loss 0.5 ||pred - obs||^2 against the record of
  the true model, gradient w.r.t. the smoothed model by boundary saving
  (fp32 strips: 42 GB per shot at nt = 2000).

``--workload acoustic2d`` (one step = one gradient over the whole shot batch):
  grid 384 x 1152 at dh = 8 m, dt = 0.8 ms, nt = 3000, 8th order in space,
  CPML 50 cells, seeded sinuous-layer vp (1500-4500 m/s, +-5% smoothed noise),
  32 surface shots (Ricker 8 Hz, delay 1.5/f) with uneven receiver spreads
  of 150-500 receivers padded to the largest count with a residual mask;
  loss 0.5 ||mask (pred - obs)||^2 against the record of the true model,
  gradient w.r.t. the smoothed starting model by boundary saving.

``--workload elastic2d``:
  grid 512 x 1024 at dh = 5 m, dt = 0.5 ms, nt = 4000, 8th order, cpmls CPML
  50 cells, seeded layered vp (1800-4500 m/s), vs = vp / 1.8 +-10%, Gardner
  rho = 310 vp^0.25; 16 shots of a vertical-force (vz) Ricker at 12 Hz, 400
  receivers each recording vx and vz; gradients w.r.t. vp, vs and rho by
  boundary saving.

``--workload acoustic2d_small``: the 96 x 128 two-layer case (nt 800, one
  shot and receiver) at order 8 with boundary saving; the reference's README
  example runs it at its default order 4.

Prints ONE JSON line: shot-gradients/s (value = shots x steps / time), the
roofline of the dominant kernel (algorithmic bytes per launch over the mean
launch duration, measured with HIP events on the launch stream during the
timed steps), and the CPU baseline: the reference's own host core
(``oracle/_ref``, built from its sources) when present, else the oracle, on
a bounded sample of one shot on this machine's host cores.
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

WORKLOADS = {
    "acoustic2d": dict(name="acoustic2d_384x1152_nt3000_32shots_bs", shape=(384, 1152), dh=8.0, dt=8e-4,
                       nt=3000, order=8, abcn=50, nshots=32, f0=8.0, seed=2, per_launch=32),
    "acoustic3d": dict(name="acoustic3d_256x512x512_nt2000_8shots_bs", shape=(256, 512, 512), dh=10.0,
                       dt=1e-3, nt=2000, order=8, abcn=50, nshots=8, f0=8.0, seed=4, per_launch=2,
                       step="launch"),
    # the reference's CPU-runnable case (its README example, run here at order 8 where the README uses the
    # default order 4): two-layer vp (1500 over 2500 m/s), one source at (x 32, z 48), one receiver at
    # (x 96, z 48), gradient of 0.5 ||pred - obs||^2 from a constant 1500 m/s start
    "acoustic2d_small": dict(name="acoustic2d_96x128_nt800_1shot_bs", shape=(96, 128), dh=10.0, dt=2e-3, nt=800,
                             order=8, abcn=50, nshots=1, f0=10.0, delay=0.14, seed=0, per_launch=1,
                             two_layer=True, src=(32, 48), rec=(96, 48),
                             note="the reference README's case at order 8 with boundary saving; the README "
                                  "runs Acoustic() at its default order 4 without checkpointing"),
    "elastic2d": dict(name="elastic2d_512x1024_nt4000_16shots_bs", shape=(512, 1024), dh=5.0, dt=5e-4,
                      nt=4000, order=8, abcn=50, nshots=16, f0=12.0, seed=6, per_launch=16, nrec=400),
}

KERNEL_NAMES = {0: "fwd (forward step + strip save)", 1: "adj (exact adjoint step)",
                2: "recon (BS reconstruction + fused imaging)"}
KERNEL_NAMES_E2D = {0: "fwd (velocity + stress step)", 1: "adj (stress/velocity adjoint prepare + apply)",
                    2: "image (vp/vs/rho imaging)"}


# ----------------------------------------------------------------- problems
def memory_strategy(spec, elastic):
    """'bs' | 'bs_bf16' | 'ckpt:<chunks>' | 'binomial:<slots>' | 'full' -> an options object
    (None: the workload's default, boundary saving)."""
    from sweep_amd.options import BoundarySaving, Ckpt, Full
    if spec is None:
        return BoundarySaving()
    kind, _, arg = spec.partition(":")
    if kind == "bs":
        return BoundarySaving()
    if kind == "bs_bf16":
        return BoundarySaving(storage_dtype="bf16")
    if kind == "full":
        return Full()
    if kind == "ckpt":
        return Ckpt(chunks=int(arg or 100))
    if kind == "binomial":
        return Ckpt(mode="recursive", count=int(arg or 100))
    raise SystemExit(f"unknown --memory {spec!r}")


def two_layer_case(W):
    """(vp_true, vp_init, sources (1, 2), receivers (1, 1, 2)) of the two-layer workload."""
    nz, nx = W["shape"]
    vp_true = np.full((nz, nx), 1500.0, np.float32)
    vp_true[nz // 2:] = 2500.0
    return (vp_true, np.full((nz, nx), 1500.0, np.float32), np.array([W["src"]], np.int64),
            np.array([[W["rec"]]], np.int64))


def make_problem(W, dev, nt, memory=None):
    """(solver, vp_true, vp_init, sources, receivers, mask or None, wavelet)."""
    from sweep_amd import synthetic
    from sweep_amd.equations import Acoustic, Acoustic3D
    from sweep_amd.options import BoundarySaving
    from sweep_amd.propagator import PropTorch
    t = np.arange(nt) * W["dt"]
    wav = synthetic.ricker(t - W.get("delay", 1.5 / W["f0"]), W["f0"])
    if W.get("two_layer"):
        vp_true, vp_init, src, rec = two_layer_case(W)
        solver = PropTorch(Acoustic(spatial_order=W["order"], device=dev), shape=W["shape"], dh=W["dh"],
                           dt=W["dt"], device=dev, abcn=W["abcn"], memory=memory_strategy(memory, False))
        return solver, [vp_true], [vp_init], src, rec, None, wav
    if W["name"].startswith("elastic"):
        from sweep_amd.equations import Elastic
        from sweep_amd.options import Ckpt
        vp_true = synthetic.layered_2d(W["shape"], seed=W["seed"], vmin=1800.0, vmax=4500.0)
        rng = np.random.default_rng(W["seed"] + 7)
        vs_true = (vp_true / 1.8 * (1.0 + 0.1 * rng.uniform(-1, 1, W["shape"]))).astype(np.float32)
        rho_true = (310.0 * vp_true.astype(np.float64) ** 0.25).astype(np.float32)
        src, rec, mask = synthetic.surface_acquisition(W["shape"], W["nshots"], seed=W["seed"],
                                                       nrec_range=(W["nrec"], W["nrec"]))
        solver = PropTorch(Elastic(spatial_order=W["order"], device=dev), shape=W["shape"], dh=W["dh"], dt=W["dt"],
                           device=dev, abcn=W["abcn"], memory=memory_strategy(memory, True))
        true = [vp_true, vs_true, rho_true]
        init = [synthetic.smoothed(m) for m in true]
        return solver, true, init, src, rec, None, wav
    if len(W["shape"]) == 2:
        vp_true = synthetic.layered_2d(W["shape"], seed=W["seed"])
        src, rec, mask = synthetic.surface_acquisition(W["shape"], W["nshots"], seed=W["seed"])
        eq = Acoustic(spatial_order=W["order"], device=dev)
    else:
        vp_true = synthetic.dipping_3d(W["shape"], seed=W["seed"])
        src, rec = synthetic.patch_acquisition_3d(W["shape"], W["nshots"])
        mask = None
        eq = Acoustic3D(spatial_order=W["order"], device=dev)
    vp_init = synthetic.smoothed(vp_true)
    solver = PropTorch(eq, shape=W["shape"], dh=W["dh"], dt=W["dt"], device=dev, abcn=W["abcn"],
                       memory=memory_strategy(memory, False))
    return solver, [vp_true], [vp_init], src, rec, mask, wav


def algorithmic_bytes(B, grid, M, pad, W):
    """Bytes per launch of the three stencil kernels under the counting rule:
    4 B per distinct full-grid float read or written per cell, perfect halo
    reuse; +16 B per cell per axis inside that axis' CPML compute band (psi,
    zeta read, next psi, zeta written); +4 B per saved strip cell; +12 B per
    imaged cell (lambda read, gradient read + write).  ``grid`` is the
    runtime grid, (nz, nx) or (nz, ny, nx)."""
    ndim = len(grid)
    comp = [n - 2 * M for n in grid]
    cells = int(np.prod(comp))
    phys = [n - 2 * pad - 2 * M for n in grid]

    def band_count(n, lo_w, hi_w):
        p = np.arange(M, n - M)
        return int(((p < lo_w) | (p >= n - hi_w)).sum())

    # per-axis band cells: the 2-D forward/adjoint ring is pad + 2M wide;
    # the 3-D forward band is pad + M + 1, the 3-D adjoint band pad + 2M + 1
    if ndim == 2:
        fb = [band_count(n, pad + 2 * M, pad + 2 * M) for n in grid]
        ab = fb
    else:
        fb = [band_count(n, pad + M + 1, pad + M + 1) for n in grid]
        ab = [band_count(n, pad + 2 * M + 1, pad + 2 * M + 1) for n in grid]
    fband = sum(fb[a] * cells // comp[a] for a in range(ndim))
    aband = sum(ab[a] * cells // comp[a] for a in range(ndim))
    if ndim == 2:
        strip = 2 * W * phys[1] + 2 * W * phys[0]
    else:
        nzp, nyp, nxp = phys
        strip = 2 * W * (nyp * nxp + nzp * nxp + nzp * nyp)
    lo = pad + 2 * M + 1
    rcells = int(np.prod([n - 2 * lo for n in grid]))
    img = int(np.prod([p - 2 * W for p in phys]))
    return {0: B * (16 * cells + 16 * fband + 4 * strip), 1: B * (16 * cells + 16 * aband),
            2: B * (16 * rcells + 12 * img)}


def elastic_bytes(B, grid, M, pad):
    """Bytes per launch class of the elastic kernels under the same rule.
    Forward: velocity {sxx, szz, sxz, rho} -> {vx, vz} (RMW) = 32 B, stress
    {vx, vz, lambda, mu} -> {sxx, szz, sxz} (RMW) = 40 B per computed cell,
    +64 B per band cell (8 memory variables read and written).  Adjoint:
    stress prepare 36 B, stress apply 32 B, velocity prepare 28 B, velocity
    apply 40 B per cell, +64 B per band cell.  Imaging: 72 B per physical
    cell (v(it), v(it+1), the adjoint fields, vp / vs / rho, three gradients
    read and written)."""
    nz, nx = grid
    cells = (nz - 2 * M) * (nx - 2 * M)
    allc = nz * nx
    inner = (nz - 2 * pad - 2 * M) * (nx - 2 * pad - 2 * M)
    band = cells - inner
    phys = inner
    return {0: B * (72 * cells + 64 * band), 1: B * (36 * allc + 28 * allc + 72 * cells + 64 * (allc - inner)),
            2: B * 72 * phys}


# ----------------------------------------------------------------- CPU leg
def cpu_baseline(W, sample_nt):
    """The oracle (C, OpenMP) on one shot of the same grid for sample_nt steps
    of forward + boundary-saving backward, extrapolated to the full job."""
    from oracle import oracle
    from sweep_amd import synthetic
    t = np.arange(sample_nt) * W["dt"]
    wav = np.ascontiguousarray(synthetic.ricker(t - W.get("delay", 1.5 / W["f0"]), W["f0"])[None, None, :])
    if W["name"].startswith("elastic"):
        P = oracle.E2DProblem(W["order"], W["shape"], W["dh"], W["dt"], sample_nt, abcn=W["abcn"], B=1)
        vp = synthetic.smoothed(synthetic.layered_2d(W["shape"], seed=W["seed"], vmin=1800.0, vmax=4500.0))
        src, rec, _ = synthetic.surface_acquisition(W["shape"], 1, seed=W["seed"], nrec_range=(W["nrec"], W["nrec"]))
        sl, rl = P.shift(src[:, None, :]), P.shift(rec)
        res = np.random.default_rng(0).standard_normal((2, 1, rl.shape[1], sample_nt)).astype(np.float32)
        models = [P.pad_model(vp), P.pad_model(vp / 1.8), P.pad_model(310.0 * vp ** 0.25)]
        t0 = time.perf_counter()
        P.gradient(*models, wav, sl, rl, res)
        dt = time.perf_counter() - t0
        scale = W["nt"] / sample_nt
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        return {"value": 1.0 / (dt * scale), "unit": "shot-gradients/s", "cores": cores, "kind": "port",
                "sample": f"1 shot, {sample_nt} of {W['nt']} steps of forward + full-storage backward on the full "
                          f"{P.nz}x{P.nx} runtime grid ({dt:.1f} s), extrapolated x{scale:g}"}
    if W.get("two_layer"):
        P = oracle.A2DProblem(W["order"], W["shape"], W["dh"], W["dt"], sample_nt, abcn=W["abcn"], B=1)
        _, vp0, src, rec = two_layer_case(W)
        vp = P.pad_model(vp0)
        grid = f"{P.nz}x{P.nx}"
    elif len(W["shape"]) == 2:
        P = oracle.A2DProblem(W["order"], W["shape"], W["dh"], W["dt"], sample_nt, abcn=W["abcn"], B=1)
        vp = P.pad_model(synthetic.smoothed(synthetic.layered_2d(W["shape"], seed=W["seed"])))
        src, rec, _ = synthetic.surface_acquisition(W["shape"], 1, seed=W["seed"], nrec_range=(300, 300))
        grid = f"{P.nz}x{P.nx}"
    else:
        P = oracle.A3DProblem(W["order"], W["shape"], W["dh"], W["dt"], sample_nt, abcn=W["abcn"], B=1)
        vp = P.pad_model(synthetic.smoothed(synthetic.dipping_3d(W["shape"], seed=W["seed"])))
        src, rec = synthetic.patch_acquisition_3d(W["shape"], 1)
        grid = f"{P.nz}x{P.ny}x{P.nx}"
    sl, rl = P.shift(src[:, None, :]), P.shift(rec)
    res = np.random.default_rng(0).standard_normal((1, rl.shape[1], sample_nt)).astype(np.float32)
    t0 = time.perf_counter()
    P.gradient(1, vp, wav, sl, rl, res)
    dt = time.perf_counter() - t0
    scale = W["nt"] / sample_nt
    per_grad = dt * scale
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1.0 / per_grad, "unit": "shot-gradients/s", "cores": cores, "kind": "port",
            "sample": f"1 shot, {sample_nt} of {W['nt']} steps of forward + boundary-saving backward "
                      f"on the full {grid} runtime grid ({dt:.1f} s), extrapolated x{scale:g}"}


# Steps of one shot the reference host core runs for the CPU baseline (a
# bounded sample of 10-30 s on the host cores, extrapolated to the job).
REF_SAMPLE_STEPS = {"acoustic3d": 8, "acoustic2d": 1000, "elastic2d": 300, "acoustic2d_small": None}


# ----------------------------------------------------------------- traffic
# HBM bytes per launch of a kernel class, from the two rocprofv3 PMC passes
# recorded in profiles/ (FETCH_SIZE scaled by 2.000 and WRITE_SIZE by 1.000,
# the factors a 1 GiB device copy reads in the same runs), keyed by workload,
# memory strategy and shots per launch at the default tile settings.
# Counters cannot be read inside this process, so the figure is the recorded
# one; any other configuration (another memory mode, shots per launch,
# --autotune, SWP_TZ* tile overrides) reports None.
MEASURED_TRAFFIC = {
    # profiles/acoustic3d_256x512x512_2per_launch_bs.md: the eight region launches of one step
    ("acoustic3d", "bs", 2): {0: 8.812e9, 1: 10.325e9, 2: 4.770e9},
    # profiles/acoustic2d_384x1152_32shots_bs.md: forward, adjoint (band + interior) and reconstruction
    ("acoustic2d", "bs", 32): {0: 462.7e6, 1: 524.3e6, 2: 433.6e6},
    # profiles/elastic2d_512x1024_16shots_bs.md: velocity + stress; prepare + apply; imaging
    ("elastic2d", "bs", 16): {0: 982.9e6, 1: 2677.6e6, 2: 875.9e6},
}
TRAFFIC_SOURCE = {
    ("acoustic3d", "bs", 2): "profiles/acoustic3d_256x512x512_2per_launch_bs.md",
    ("acoustic2d", "bs", 32): "profiles/acoustic2d_384x1152_32shots_bs.md",
    ("elastic2d", "bs", 16): "profiles/elastic2d_512x1024_16shots_bs.md",
}
TILE_ENV = ("SWP_TZ", "SWP_TZ_BAND", "SWP_TZ_RECON", "SWP_TZ3", "SWP_REGION_WAVES", "SWP_KERNELS", "SWP_SHELL3")


def measured_traffic(workload, memory, cls, per_launch, autotuned):
    """(bytes, source) recorded for this exact configuration, else (None, why)."""
    key = (workload, memory or "bs", per_launch)
    if autotuned or any(os.environ.get(v) for v in TILE_ENV):
        return None, "tile settings differ from the recorded profile"
    if key not in MEASURED_TRAFFIC or cls not in MEASURED_TRAFFIC[key]:
        return None, "no PMC profile recorded for this configuration"
    return MEASURED_TRAFFIC[key][cls], "recorded: " + TRAFFIC_SOURCE[key]


# ----------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="acoustic3d")
    ap.add_argument("--timing-stride", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-port", action="store_true",
                    help="time the oracle (the C restatement) as the CPU baseline even where the reference's "
                         "host core is built")
    ap.add_argument("--cpu-sample-nt", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None, help="override nt (profiling runs only)")
    ap.add_argument("--shots-per-launch", type=int, default=None)
    ap.add_argument("--memory", default=None,
                    help="gradient memory strategy: bs | bs_bf16 | full | ckpt:<chunks> | binomial:<slots> "
                         "(default: the workload's)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick shots per launch and the z-march segment heights by timing short probes first "
                         "(untimed); the choice is reported under config.autotuned")
    ap.add_argument("--graphs", choices=["on", "off"], default="on",
                    help="replay each launch's time loop from a captured HIP graph (on) or launch step by step")
    ap.add_argument("--pmc-launch", action="store_true",
                    help="counter runs: a 1 GiB device copy (the calibration of FETCH_SIZE / WRITE_SIZE), "
                         "then one untimed gradient launch; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from sweep_amd import _core
    from sweep_amd.loss import l2_misfit
    W = WORKLOADS[args.workload]
    _core.set_graphs(1 if args.graphs == "on" else 0)
    nt = args.nt or W["nt"]
    per_launch = args.shots_per_launch or W["per_launch"]
    dev = torch.device("cuda:0")
    solver, vp_true, vp_init, src, rec, mask, wav = make_problem(W, dev, nt, args.memory)
    nshots = src.shape[0]
    if args.pmc_launch:
        ca = torch.ones(1 << 28, device=dev)
        cb = torch.empty_like(ca)
        cb.copy_(ca)
        torch.cuda.synchronize()
        del ca, cb
        ms = [torch.tensor(m, device=dev, requires_grad=True) for m in vp_init]
        pred = solver(wav, src[:per_launch], rec[:per_launch], models=ms)
        (0.5 * pred.pow(2).sum()).backward()
        torch.cuda.synchronize()
        print(f"[bench] {W['name']} nt={nt}: one launch of {min(per_launch, nshots)} shots done", file=sys.stderr,
              flush=True)
        return
    autotuned = None
    if args.autotune:
        from sweep_amd.autotune import tune
        print(f"[bench] {W['name']}: autotuning", file=sys.stderr, flush=True)
        res = tune(solver, wav, src, rec, [torch.tensor(m, device=dev) for m in vp_init], nshots, full_nt=nt)
        per_launch = res.shots_per_launch
        autotuned = {"shots_per_launch": per_launch, **res.knobs,
                     "probe_us_per_shot_step": round(res.seconds_per_shot_step * 1e6, 3)}
        print(f"[bench] autotuned {autotuned}", file=sys.stderr, flush=True)
    batches = [(b, min(b + per_launch, nshots)) for b in range(0, nshots, per_launch)]
    print(f"[bench] {W['name']} nt={nt}: building obs", file=sys.stderr, flush=True)
    vt = [torch.tensor(m, device=dev) for m in vp_true]
    with torch.no_grad():
        obs = [solver(wav, src[a:b], rec[a:b], models=vt) for a, b in batches]
    del vt
    masks = [None if mask is None else torch.tensor(mask[a:b], device=dev) for a, b in batches]
    m0 = [torch.tensor(m, device=dev) for m in vp_init]

    per_step_launch = W.get("step") == "launch"
    launches = [0]

    def step():
        """One gradient over all batches, or (per_step_launch) one batch in turn."""
        ms = [m.clone().requires_grad_(True) for m in m0]
        todo = [launches[0] % len(batches)] if per_step_launch else range(len(batches))
        for k in todo:
            a, b = batches[k]
            pred = solver(wav, src[a:b], rec[a:b], models=ms)
            loss = l2_misfit(pred, obs[k], masks[k])
            loss.backward()
            launches[0] += 1
        return ms[0].grad

    for k in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        g = step()
        print(f"[bench] step {k + 1}/{args.steps} launched", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    graphs = _core.graph_stats()
    assert torch.isfinite(g).all(), "non-finite gradient"
    # kernel durations: one more step, launched eagerly with HIP events around
    # one launch in every timing_stride of each kernel class (a launch's
    # duration does not depend on how it was enqueued)
    _core.set_timing(args.timing_stride)
    step()
    torch.cuda.synchronize()
    stats = {c: _core.kernel_stats(c) for c in range(4)}
    _core.set_timing(0)
    print(f"[bench] kernel timing step done; graphs {graphs}", file=sys.stderr, flush=True)

    P = solver._last_call.problem
    grid = (P.nz, P.nx) if P.ndim == 2 else (P.nz, P.ny, P.nx)
    tape_bytes = int(P.tape_bytes)
    elastic = W["name"].startswith("elastic")
    names = KERNEL_NAMES_E2D if elastic else KERNEL_NAMES
    abytes = elastic_bytes(P.B, grid, P.M, W["abcn"]) if elastic else algorithmic_bytes(P.B, grid, P.M, W["abcn"],
                                                                                        P.M + 1)
    per = {}
    for c in (0, 1, 2):
        ms, n = stats[c]
        if n:
            per[c] = dict(avg_ms=ms / n, total_est_ms=ms * args.timing_stride,
                          gbs=abytes[c] / (ms / n * 1e-3) / 1e9)
    dom = max(per, key=lambda c: per[c]["total_est_ms"])
    ach = per[dom]["gbs"]
    traffic, traffic_src = measured_traffic(args.workload, args.memory, dom, per_launch, autotuned)
    shots_per_step = per_launch if per_step_launch else nshots
    value = shots_per_step * args.steps / elapsed
    out = {
        "metric": "shot-gradients/s fwd+adj", "value": value, "unit": "shot-gradients/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": W["name"] if args.nt is None else f"{W['name']}_nt{nt}",
                   "grid": list(W["shape"]), "runtime_grid": list(grid), "nt": nt, "shots": nshots,
                   "shots_per_launch": per_launch, "order": W["order"],
                   "step": (f"one launch of {per_launch} of the {nshots} shots, taken in turn" if per_step_launch
                            else f"one gradient over all {nshots} shots"),
                   "memory": args.memory or "bs",
                   "tape_bytes_per_launch": tape_bytes,
                   "graphs": args.graphs,
                   **({"note": W["note"]} if "note" in W else {}),
                   **({"autotuned": autotuned} if autotuned else {})},
        "roofline": {"bound": "hbm", "kernel": names[dom], "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": int(abytes[dom]),
                     "avg_launch_ms": per[dom]["avg_ms"],
                     "kernels": {names[c]: {"avg_launch_ms": per[c]["avg_ms"], "GB/s": per[c]["gbs"],
                                                   "frac": per[c]["gbs"] / HBM_PEAK_GBS} for c in per}},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        from oracle import refcore
        if refcore.available() and not args.cpu_port:
            # the reference's own CPU path: its host core built from its sources
            from oracle import time_ref
            steps = args.cpu_sample_nt or REF_SAMPLE_STEPS[args.workload]
            print(f"[bench] timing the reference host core on {steps} steps of one shot", file=sys.stderr,
                  flush=True)
            ref = time_ref.time_sample(args.workload, steps, order=W["order"])
            out["cpu_baseline"] = {k: ref[k] for k in ("value", "unit", "cores", "kind", "sample")}
        else:
            print("[bench] timing the CPU oracle sample", file=sys.stderr, flush=True)
            sample = args.cpu_sample_nt or (8 if len(W["shape"]) == 3 else (3000 if elastic else W["nt"]))
            out["cpu_baseline"] = cpu_baseline(W, sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
