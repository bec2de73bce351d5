"""sweep_call (the reference's C ABI, version 5) against swp_call on the GPU.

Each case fills a ForwardInputCore and a BackwardInputCore the way the
reference's wrapper does (propagator/_c.py:410-540, 620-760): models on the
padded runtime grid, the wavelet as (B, nsrc, nt), int32 source/receiver
locations, the CPML profiles, host FD coefficients, the memory-strategy
flags, grads_out = [grad_wavelet when the equation returns it, one per
model] -- with this core's workspace in forward_workspace[0] /
adjoint_workspace[0] and its tape where the reference keeps the memory
mode's state (u_allt_out / u_forward, boundary_gpu[0], checkpoints[0]),
sized by sweep_query_bytes.  The record and every gradient must equal, bit for bit,
what PropTorch (which calls swp_call) returns for the same inputs.
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


def _case(kind, dev):
    from sweep_amd import synthetic
    from sweep_amd.equations import Acoustic, Acoustic3D, Elastic
    from sweep_amd.options import BoundarySaving, Ckpt, Full
    rng = np.random.default_rng(3)
    if kind.startswith("a2d"):
        shape, nt, dt, dh, B = (48, 64), 300, 1e-3, (10.0, 12.0), 2
        vp = synthetic.layered_2d(shape, seed=4, vmin=1500.0, vmax=3000.0)
        src, rec, _ = synthetic.surface_acquisition(shape, B, seed=4, nrec_range=(20, 20), depth=3)
        eq, models, nf = Acoustic(spatial_order=8, device=dev), [vp], 1
    elif kind.startswith("a3d"):
        shape, nt, dt, dh, B = (20, 18, 22), 120, 1e-3, (10.0, 11.0, 12.0), 2
        vp = (1800.0 + 500.0 * rng.random(shape)).astype(np.float32)
        src = np.stack([rng.integers(3, 19, B), rng.integers(3, 15, B), np.full(B, 3)], -1)
        rec = np.stack([rng.integers(0, 22, (B, 12)), rng.integers(0, 18, (B, 12)), np.full((B, 12), 2)], -1)
        eq, models, nf = Acoustic3D(spatial_order=8, device=dev), [vp], 1
    else:
        shape, nt, dt, dh, B = (48, 64), 300, 5e-4, (5.0, 6.0), 2
        vp = synthetic.smoothed(synthetic.layered_2d(shape, seed=5, vmin=2000.0, vmax=4000.0), sigma=3.0)
        src, rec, _ = synthetic.surface_acquisition(shape, B, seed=5, nrec_range=(16, 16), depth=3)
        eq = Elastic(spatial_order=8, device=dev)
        models = [vp, (vp / 1.8).astype(np.float32), (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)]
        nf = 2
    memory = {"bs": BoundarySaving(), "full": Full(), "ckpt": Ckpt(chunks=6)}[kind.split("_")[1]]
    wav = np.stack([synthetic.ricker(np.arange(nt) * dt - 0.06, 15.0) * (1 + 0.1 * b) for b in range(B)])
    res = rng.standard_normal((B, nt, rec.shape[1], nf)).astype(np.float32)
    return dict(shape=shape, nt=nt, dt=dt, dh=dh, B=B, eq=eq, memory=memory, models=models, src=src, rec=rec,
                wav=wav.astype(np.float32), res=res, nf=nf)


def _via_proptorch(c, dev):
    from sweep_amd.propagator import PropTorch
    solver = PropTorch(c["eq"], shape=c["shape"], dh=c["dh"], dt=c["dt"], device=dev, abcn=8, memory=c["memory"])
    ms = [torch.tensor(m, device=dev, requires_grad=True) for m in c["models"]]
    elastic = c["nf"] == 2
    w = torch.tensor(c["wav"], device=dev, requires_grad=not elastic)
    pred = solver(w, c["src"], c["rec"], models=ms)
    (pred * torch.tensor(c["res"], device=dev)).sum().backward()
    torch.cuda.synchronize()
    out = {"rec": pred.detach().cpu().numpy(), "grads": [m.grad.cpu().numpy() for m in ms]}
    if not elastic:
        out["grad_w"] = w.grad.cpu().numpy()
    return solver, out


def _via_sweep_call(solver, c, dev):
    from sweep_amd import abi5, geometry, records
    call = solver._last_call
    P = call.problem
    src_t, rec_t, prof_t, lap, grad = call.keep[:5]
    ndim = solver.ndim
    pad = geometry.torch_pad_order(solver.pad, ndim, solver.equation.M)
    padded = [geometry.EdgePadding.apply(torch.tensor(m, device=dev), pad).contiguous() for m in c["models"]]
    w3 = torch.tensor(c["wav"], device=dev)[:, None, :].contiguous()
    B, nt, nf = c["B"], c["nt"], c["nf"]
    nrec = c["rec"].shape[1]
    keep = []

    def lst(ts):
        bl, arr = abi5.buf_list(ts)
        keep.append(arr)
        return bl

    def span(kind, vals):
        s, arr = (abi5.int_span if kind == "i" else abi5.float_span)(vals)
        keep.append(arr)
        return s

    lap_t, grad_t = torch.from_numpy(np.ascontiguousarray(lap)), torch.from_numpy(np.ascontiguousarray(grad))
    sf = [P.src_fields[k] for k in range(P.nsrc_fields)]
    rf = [P.rec_fields[k] for k in range(P.nrec_fields)]
    spacing = list(reversed(solver.dh))   # Cartesian (x[, y], z), propagator/_c.py:1129-1132
    fin = abi5.ForwardInputCore()
    fin.models = lst(padded)
    fin.source = abi5.buf_of(w3)
    fin.lap_coes, fin.grad_coes = abi5.buf_of(lap_t), abi5.buf_of(grad_t)
    fin.M, fin.abcn = solver.equation.M, solver.abcn
    fin.sources_loc, fin.receivers_loc = abi5.buf_of(src_t), abi5.buf_of(rec_t)
    fin.source_field_indices, fin.receiver_field_indices = span("i", sf), span("i", rf)
    fin.pml_vals = lst(list(prof_t))
    fin.nt, fin.dt, fin.spacing = nt, c["dt"], span("f", spacing)
    mem = int(call.memory)
    from sweep_amd import _core
    fin.use_boundary_saving = mem == _core.MEM_BOUNDARY
    fin.use_checkpoint = mem == _core.MEM_CKPT
    fin.save_all_wavefields = mem == _core.MEM_FULL
    fin.checkpoint_interval = int(P.ckpt_interval)
    names = call.names
    ws_f, tp = abi5.query_bytes(names[0], fin)
    ws_f_t = torch.empty(ws_f, dtype=torch.uint8, device=dev)
    tape = torch.empty(tp, dtype=torch.uint8, device=dev)
    fin.forward_workspace = lst([ws_f_t])
    # the tape where the reference keeps the mode's state
    if mem == _core.MEM_FULL:
        fin.u_allt_out = abi5.buf_of(tape)
    elif mem == _core.MEM_BOUNDARY:
        fin.boundary_gpu = lst([tape])
    else:
        fin.checkpoints = lst([tape])
    raw = torch.zeros((nf, B, nrec, nt) if nf > 1 else (B, nrec, nt), device=dev)
    fin.record_out = abi5.buf_of(raw)
    stream = torch.cuda.current_stream(dev).cuda_stream
    abi5.call(names[0], fin, stream)

    bk = {_core.MEM_FULL: 1, _core.MEM_BOUNDARY: 2, _core.MEM_CKPT: 3}[mem]
    bin_ = abi5.BackwardInputCore()
    bin_.models = fin.models
    bin_.forward_source = fin.source
    bin_.lap_coes, bin_.grad_coes = fin.lap_coes, fin.grad_coes
    bin_.M, bin_.abcn = fin.M, fin.abcn
    bin_.forward_sources_loc, bin_.adjoint_sources_loc = fin.sources_loc, fin.receivers_loc
    bin_.source_field_indices, bin_.receiver_field_indices = fin.source_field_indices, fin.receiver_field_indices
    bin_.pml_vals = fin.pml_vals
    bin_.nt, bin_.dt, bin_.spacing = nt, c["dt"], fin.spacing
    bin_.checkpoint_interval = fin.checkpoint_interval
    bin_.compute_illumination = False
    adj = records.to_raw(torch.tensor(c["res"], device=dev), 4 if nf > 1 else 3)
    bin_.adjoint_source = abi5.buf_of(adj)
    grid = (B,) + tuple(padded[0].shape)
    gm = [torch.zeros(grid, device=dev) for _ in padded]
    gw = torch.zeros_like(w3) if nf == 1 else None
    bin_.grads_out = lst(([gw] if gw is not None else []) + gm)
    ws_b, _ = abi5.query_bytes(names[bk], bin_)
    ws_b_t = torch.empty(ws_b, dtype=torch.uint8, device=dev)
    bin_.adjoint_workspace = lst([ws_b_t])
    bin_.u_forward, bin_.boundary_gpu, bin_.checkpoints = fin.u_allt_out, fin.boundary_gpu, fin.checkpoints
    abi5.call(names[bk], bin_, stream)
    torch.cuda.synchronize()
    crop = [geometry.crop_physical(g.sum(0), solver.pad, solver.equation.M, ndim) for g in gm]
    out = {"rec": records.to_canonical(raw).cpu().numpy(), "grads": [g.cpu().numpy() for g in crop]}
    if gw is not None:
        out["grad_w"] = gw[:, 0].cpu().numpy()
    return out


@pytest.mark.parametrize("kind", ["a2d_bs", "a2d_full", "a3d_bs", "a3d_ckpt", "e2d_bs"])
def test_sweep_call_matches_swp_call(kind):
    dev = _dev()
    print(f"[abi5] case={kind}", flush=True)
    c = _case(kind, dev)
    solver, ref = _via_proptorch(c, dev)
    got = _via_sweep_call(solver, c, dev)
    assert np.abs(ref["rec"]).max() > 0
    assert np.array_equal(got["rec"], ref["rec"])
    for k, (a, b) in enumerate(zip(got["grads"], ref["grads"])):
        assert np.abs(b).max() > 0, k
        assert np.array_equal(a, b), k
    if "grad_w" in ref:
        assert np.array_equal(got["grad_w"], ref["grad_w"])
