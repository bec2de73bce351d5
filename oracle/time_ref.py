"""ORACLE -- test infrastructure only: times the reference's CPU path.

The reference's own CPU host core (its ``impl='c'`` on the CPU: the same
kernel sources compiled by g++ with OpenMP, ``backend/c/hostcore.py:64-82``),
built from its sources into ``oracle/_ref/`` by ``oracle/build_ref.sh`` and
driven through ``oracle/ref_harness.cpp``, on one shot of each benchmark
workload: forward + boundary-saving backward of a seeded residual.

* The 96 x 128 case (order 4 as in the reference's README example, nt 800)
  runs whole.
* The full-size workloads run ``--steps`` (default 50) forward and 50
  backward steps of one shot on the exact runtime grid after one warm-up
  run of 4 steps, and the per-step time is extrapolated to the workload's nt
  and shot count (stated in the output).

``python -m oracle.time_ref [--workloads ...] [--steps K]`` prints one JSON
line per workload (ms per forward+backward step and per forward step,
shot-gradients/s, the extrapolation, OMP_NUM_THREADS).  ``bench.py`` imports :func:`time_sample`
for its ``cpu_baseline`` leg (kind "reference").
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# name -> (equation, physical shape, dh, dt, nt, order, shots, f0)
WORKLOADS = {
    "acoustic2d_small": ("acoustic", (96, 128), 10.0, 2e-3, 800, 4, 1, 10.0),
    "acoustic2d": ("acoustic", (384, 1152), 8.0, 8e-4, 3000, 8, 32, 8.0),
    "elastic2d": ("elastic", (512, 1024), 5.0, 5e-4, 4000, 8, 16, 12.0),
    "acoustic3d": ("acoustic", (256, 512, 512), 10.0, 1e-3, 2000, 8, 8, 8.0),
}


def _inputs(name, nt, order=None):
    """(problem, padded models, wavelet (1, 1, nt), sources, receivers, residual)."""
    from oracle import oracle
    from sweep_amd import synthetic
    eq, shape, dh, dt, _, order0, _, f0 = WORKLOADS[name]
    order = order or order0
    wav = synthetic.ricker(np.arange(nt) * dt - 1.5 / f0, f0)[None, None, :]
    rng = np.random.default_rng(0)
    if name == "acoustic2d_small":
        P = oracle.A2DProblem(order, shape, dh, dt, nt, abcn=50, B=1)
        vp = np.full(shape, 1500.0, np.float32)
        src, rec = np.array([[[32, 48]]]), np.array([[[96, 48]]])
        wav = synthetic.ricker(np.arange(nt) * dt - 0.14, f0)[None, None, :]
        models = [P.pad_model(vp)]
    elif eq == "elastic":
        P = oracle.E2DProblem(order, shape, dh, dt, nt, abcn=50, B=1)
        vp = synthetic.smoothed(synthetic.layered_2d(shape, seed=6, vmin=1800.0, vmax=4500.0))
        src2, rec2, _ = synthetic.surface_acquisition(shape, 1, seed=6, nrec_range=(400, 400))
        src, rec = src2[:, None, :], rec2
        models = [P.pad_model(vp), P.pad_model(vp / 1.8), P.pad_model(310.0 * vp ** 0.25)]
    elif len(shape) == 2:
        P = oracle.A2DProblem(order, shape, dh, dt, nt, abcn=50, B=1)
        vp = synthetic.smoothed(synthetic.layered_2d(shape, seed=2))
        src2, rec2, _ = synthetic.surface_acquisition(shape, 1, seed=2, nrec_range=(500, 500))
        src, rec = src2[:, None, :], rec2
        models = [P.pad_model(vp)]
    else:
        P = oracle.A3DProblem(order, shape, dh, dt, nt, abcn=50, B=1)
        vp = synthetic.smoothed(synthetic.dipping_3d(shape, seed=4))
        src3, rec3 = synthetic.patch_acquisition_3d(shape, 1)
        src, rec = src3[:, None, :], rec3
        models = [P.pad_model(vp)]
    nf = 2 if eq == "elastic" else 0
    res_shape = ((nf,) if nf else ()) + (1, rec.shape[1], nt)
    res = rng.standard_normal(res_shape).astype(np.float32)
    return P, models, np.ascontiguousarray(wav, np.float32), P.shift(src), P.shift(rec), res


def _run_once(name, nt, order=None, backward=True):
    """Seconds of one reference run: forward, plus the boundary-saving
    backward of a seeded residual when ``backward``."""
    from oracle import refcore
    P, models, wav, sl, rl, res = _inputs(name, nt, order)
    res = res if backward else None
    t0 = time.perf_counter()
    if WORKLOADS[name][0] == "elastic":
        refcore.run_elastic((P.nz, P.nx), P.M, nt, P.abcn, P.dt, P.dh_zyx, P.prof, *models, P.grad, wav, sl, rl,
                            res, bs=backward)
    else:
        refcore.run(P, models[0], wav, sl, rl, res, bs=backward)
    return time.perf_counter() - t0, P


def time_forward(name, steps=None, order=None):
    """ms per forward step of the reference host core on one shot."""
    nt = WORKLOADS[name][4]
    k = nt if steps is None else min(steps, nt)
    sec, _ = _run_once(name, k, order, backward=False)
    return sec / k * 1e3


def time_sample(name, steps=None, warm=True, order=None):
    """The reference host core on one shot of workload ``name`` (at its
    spatial order, or ``order``): the whole run when ``steps`` is None, else
    ``steps`` forward + ``steps`` backward steps extrapolated to the
    workload's nt and shots.  Returns a dict with shot-gradients/s and how it
    was measured."""
    eq, shape, dh, dt, nt, order0, shots, f0 = WORKLOADS[name]
    order = order or order0
    k = nt if steps is None else min(steps, nt)
    if warm and k != 4:
        _run_once(name, 4, order)
    sec, P = _run_once(name, k, order)
    grid = "x".join(str(n) for n in ((P.nz, P.ny, P.nx) if hasattr(P, "ny") else (P.nz, P.nx)))
    per_step = sec / k
    per_grad = per_step * nt
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"workload": name, "ms_per_step": per_step * 1e3, "value": 1.0 / per_grad, "unit": "shot-gradients/s",
            "seconds_per_job": per_grad * shots, "cores": threads, "kind": "reference",
            "order": order,
            "sample": f"the reference's host core (oracle/_ref), order {order}, 1 shot, {k} of {nt} forward + "
                      f"boundary-saving "
                      f"backward steps on the {grid} runtime grid ({sec:.1f} s), extrapolated x{nt / k:g} per shot"
                      f" and x{shots} shots for the job"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", nargs="*", default=list(WORKLOADS))
    ap.add_argument("--steps", type=int, default=50)
    args = ap.parse_args()
    from oracle import refcore
    if not refcore.available():
        raise SystemExit("oracle/_ref is not built (oracle/build_ref.sh needs the reference tree)")
    for name in args.workloads:
        steps = None if name == "acoustic2d_small" else args.steps
        out = time_sample(name, steps)
        out["fwd_ms_per_step"] = time_forward(name, steps)
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
