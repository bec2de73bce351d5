"""Time-loop graphs: a call captured into a HIP graph and replayed gives the
eager launches' results bit for bit.

Each case runs once with graphs off (the eager reference), then several
times with graphs on: the first graphed call captures the whole forward
(resp. backward) launch sequence, later identical calls replay it.  Between
the runs the test does unrelated eager torch work and synchronizes, and it
checks every replay against the eager result: record, model gradients and
wavelet gradient bit-identical.  The graph counters must show captures and
replays, so the replays really ran.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _graphs_default():
    yield
    from sweep_amd import _core
    _core.set_graphs(-1)
    _core.graph_clear()


def _case(kind, dev):
    from sweep_amd import synthetic
    from sweep_amd.equations import Acoustic, Acoustic3D, Elastic
    from sweep_amd.options import BoundarySaving, Full
    rng = np.random.default_rng(7)
    if kind in ("a2d_small", "a2d_zm_bs", "a2d_zm_full"):
        shape, nt, dt, dh = ((96, 128), 300, 1.5e-3, 10.0) if kind == "a2d_small" else ((96, 256), 250, 8e-4, 8.0)
        B = 1 if kind == "a2d_small" else 3
        vp = synthetic.layered_2d(shape, seed=3, vmin=1500.0, vmax=3000.0)
        src, rec, _ = synthetic.surface_acquisition(shape, B, seed=3, nrec_range=(20, 40), depth=10)
        mem = Full() if kind == "a2d_zm_full" else BoundarySaving()
        from sweep_amd.propagator import PropTorch
        solver = PropTorch(Acoustic(spatial_order=8, device=dev), shape=shape, dh=dh, dt=dt, device=dev,
                           abcn=20 if B > 1 else 50, memory=mem)
        models = [vp]
        nf = 1
    elif kind == "a3d_zm_bs":
        shape, nt, dt, dh = (40, 36, 96), 100, 1e-3, 10.0
        vp = (1800.0 + 600.0 * rng.random(shape)).astype(np.float32)
        B = 2
        src = np.stack([rng.integers(20, 76, B), rng.integers(8, 28, B), np.full(B, 12)], -1)
        rec = np.stack([rng.integers(0, 96, (B, 30)), rng.integers(0, 36, (B, 30)), np.full((B, 30), 2)], -1)
        from sweep_amd.propagator import PropTorch
        solver = PropTorch(Acoustic3D(spatial_order=8, device=dev), shape=shape, dh=dh, dt=dt, device=dev, abcn=10,
                           memory=BoundarySaving())
        models = [vp]
        nf = 1
    else:   # elastic, boundary saving
        shape, nt, dt, dh = (64, 96), 300, 5e-4, 5.0
        vp = synthetic.smoothed(synthetic.layered_2d(shape, seed=8, vmin=2000.0, vmax=4000.0), sigma=3.0)
        B = 2
        src, rec, _ = synthetic.surface_acquisition(shape, B, seed=8, nrec_range=(30, 30))
        from sweep_amd.propagator import PropTorch
        solver = PropTorch(Elastic(spatial_order=8, device=dev), shape=shape, dh=dh, dt=dt, device=dev, abcn=20,
                           memory=BoundarySaving())
        models = [vp, (vp / 1.8).astype(np.float32), (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)]
        nf = 2
    wav = np.stack([synthetic.ricker(np.arange(nt) * dt - 0.08, 12.0) * (1 + 0.1 * b) for b in range(B)])
    res = rng.standard_normal((B, nt, rec.shape[1], nf)).astype(np.float32)
    return solver, models, wav.astype(np.float32), src, rec, res, kind != "elastic_bs"


def _run(solver, models, wav, src, rec, res, grad_w, dev):
    ms = [torch.tensor(m, device=dev, requires_grad=True) for m in models]
    w = torch.tensor(wav, device=dev, requires_grad=grad_w)
    pred = solver(w, src, rec, models=ms)
    (pred * torch.tensor(res, device=dev)).sum().backward()
    torch.cuda.synchronize()
    out = {"rec": pred.detach().cpu().numpy()}
    for k, m in enumerate(ms):
        out[f"grad{k}"] = m.grad.cpu().numpy()
    if grad_w:
        out["grad_w"] = w.grad.cpu().numpy()
    return out


@pytest.mark.parametrize("kind", ["a2d_small", "a2d_zm_bs", "a2d_zm_full", "a3d_zm_bs", "elastic_bs"])
def test_graph_replays_match_eager(kind, monkeypatch):
    dev = _dev()
    from sweep_amd import _core
    if "zm" in kind:
        monkeypatch.setenv("SWP_KERNELS", "zmarch")
    print(f"[graphs] case={kind}", flush=True)
    solver, models, wav, src, rec, res, grad_w = _case(kind, dev)
    _core.set_graphs(0)
    ref = _run(solver, models, wav, src, rec, res, grad_w, dev)
    for name in ref:
        assert np.isfinite(ref[name]).all() and np.abs(ref[name]).max() > 0, name
    _core.set_graphs(1)
    _core.graph_clear()
    for k in range(4):
        junk = torch.randn(512, 512, device=dev)
        (junk @ junk).sum().item()
        torch.cuda.synchronize()
        got = _run(solver, models, wav, src, rec, res, grad_w, dev)
        for name in ref:
            assert np.array_equal(got[name], ref[name]), (kind, k, name)
    st = _core.graph_stats()
    print(f"[graphs] {kind} {st}", flush=True)
    assert st["captures"] >= 2 and st["replays"] >= 2, st


def test_graphs_off_runs_eagerly():
    dev = _dev()
    from sweep_amd import _core
    solver, models, wav, src, rec, res, grad_w = _case("a2d_small", dev)
    _core.set_graphs(0)
    _core.graph_clear()
    _run(solver, models, wav, src, rec, res, grad_w, dev)
    assert _core.graph_stats()["captures"] == 0


def test_long_reversals_run_eagerly():
    """A binomial reversal with one snapshot slot re-runs O(nt^2) steps; the
    core launches it step by step instead of capturing a graph of ~10^5
    nodes (only the forward, nt steps, is captured)."""
    dev = _dev()
    from sweep_amd import _core, synthetic
    from sweep_amd.equations import Acoustic
    from sweep_amd.options import Ckpt
    from sweep_amd.propagator import PropTorch
    shape, nt, dt = (40, 48), 400, 1e-3
    solver = PropTorch(Acoustic(spatial_order=8, device=dev), shape=shape, dh=10.0, dt=dt, device=dev, abcn=10,
                       memory=Ckpt(mode="recursive", count=1))
    vp = torch.full(shape, 2000.0, device=dev, requires_grad=True)
    wav = synthetic.ricker(np.arange(nt) * dt - 0.08, 15.0)
    src = np.array([[20, 3]])
    rec = np.array([[[4 + 4 * k, 2] for k in range(8)]])
    _core.set_graphs(1)
    _core.graph_clear()
    (solver(wav, src, rec, models=[vp]) ** 2).sum().backward()
    torch.cuda.synchronize()
    st = _core.graph_stats()
    print(f"[graphs] binomial, one slot: {st}", flush=True)
    assert st["captures"] == 1, st
    assert torch.isfinite(vp.grad).all() and vp.grad.abs().max() > 0
