"""Launch-shape tuning on the GPU: records are bit-identical whatever the
segment heights, 3-D gradients too; 2-D gradients may move at the float32
rounding level when the reconstruction segment height changes (measured
4e-8 relative; bar 1e-5, a tenth of the gradient parity bar).  The
autotuner returns a configuration that fits and is the fastest it timed."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from sweep_amd import _core, synthetic
from sweep_amd.autotune import tune
from sweep_amd.equations import Acoustic, Acoustic3D
from sweep_amd.options import BoundarySaving
from sweep_amd.propagator import PropTorch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _rel(a, b):
    return float((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-300))


def _grad(solver, wav, src, rec, vp):
    m = vp.clone().requires_grad_(True)
    pred = solver(wav, src, rec, models=[m])
    (0.5 * pred.pow(2).sum()).backward()
    return pred.detach(), m.grad


@pytest.fixture(autouse=True)
def _reset(monkeypatch):
    # the 2-D cases here are small enough for the per-cell kernels: force
    # the z-marching ones, whose segment heights are what is tuned
    monkeypatch.setenv("SWP_KERNELS", "zmarch")
    yield
    for k in (_core.TUNE_TZ2, _core.TUNE_TZ2_BAND, _core.TUNE_TZ2_RECON, _core.TUNE_TZ3):
        _core.set_tuning(k, 0)


def test_2d_results_do_not_depend_on_segment_heights():
    print("[autotune] 2-D segment heights", flush=True)
    shape = (96, 200)
    vp = torch.tensor(synthetic.layered_2d(shape, seed=2, vmin=1500.0, vmax=3000.0, nlayers=4), device=DEV)
    src, rec, _ = synthetic.surface_acquisition(shape, 3, seed=2, nrec_range=(50, 80))
    wav = synthetic.ricker(np.arange(400) * 1e-3 - 0.1, 12.0)
    solver = PropTorch(Acoustic(spatial_order=8, device=DEV), shape=shape, dh=10.0, dt=1e-3, device=DEV, abcn=20,
                       memory=BoundarySaving())
    r0, g0 = _grad(solver, wav, src, rec, vp)
    for tz, tzb, tzr in ((8, 4, 32), (64, 16, 8), (5, 3, 7)):
        _core.set_tuning(_core.TUNE_TZ2, tz)
        _core.set_tuning(_core.TUNE_TZ2_BAND, tzb)
        _core.set_tuning(_core.TUNE_TZ2_RECON, tzr)
        r, g = _grad(solver, wav, src, rec, vp)
        er, eg = _rel(r, r0), _rel(g, g0)
        print(f"[autotune] 2-D tz {tz} band {tzb} recon {tzr}: record {er:.2e} gradient {eg:.2e} "
              f"identical {bool(torch.equal(r, r0))}/{bool(torch.equal(g, g0))}", flush=True)
        assert er == 0.0 and eg <= 1e-5, (tz, tzb, tzr)


def test_3d_results_do_not_depend_on_segment_heights():
    print("[autotune] 3-D segment heights", flush=True)
    shape = (40, 36, 44)
    vp = torch.tensor(synthetic.dipping_3d(shape, seed=1), device=DEV)
    src, rec = synthetic.patch_acquisition_3d(shape, 2, patch=8)
    wav = synthetic.ricker(np.arange(150) * 1e-3 - 0.08, 15.0)
    solver = PropTorch(Acoustic3D(spatial_order=8, device=DEV), shape=shape, dh=10.0, dt=1e-3, device=DEV,
                       abcn=10, memory=BoundarySaving())
    r0, g0 = _grad(solver, wav, src, rec, vp)
    for tz in (8, 13, 64):
        _core.set_tuning(_core.TUNE_TZ3, tz)
        r, g = _grad(solver, wav, src, rec, vp)
        er, eg = _rel(r, r0), _rel(g, g0)
        print(f"[autotune] 3-D tz {tz}: record {er:.2e} gradient {eg:.2e} "
              f"identical {bool(torch.equal(r, r0))}/{bool(torch.equal(g, g0))}", flush=True)
        assert er == 0.0 and eg == 0.0, tz


def test_tuner_picks_the_fastest_fitting_configuration():
    print("[autotune] tuner on a small 2-D workload", flush=True)
    shape = (96, 256)
    vp = torch.tensor(synthetic.layered_2d(shape, seed=4, vmin=1500.0, vmax=3000.0, nlayers=4), device=DEV)
    src, rec, _ = synthetic.surface_acquisition(shape, 8, seed=4, nrec_range=(50, 80))
    wav = synthetic.ricker(np.arange(300) * 1e-3 - 0.1, 12.0)
    solver = PropTorch(Acoustic(spatial_order=8, device=DEV), shape=shape, dh=10.0, dt=1e-3, device=DEV, abcn=20,
                       memory=BoundarySaving())
    res = tune(solver, wav, src, rec, [vp], nshots=8, probe_nt=40, repeats=1)
    print(f"[autotune] chose {res.shots_per_launch} per launch, {res.knobs}", flush=True)
    assert res.shots_per_launch in (1, 2, 4, 8)
    assert set(res.knobs) == {"tz2", "tz2_band", "tz2_recon"}
    assert res.seconds_per_shot_step == min(t for _, _, t in res.table)
    for name, key in (("tz2", _core.TUNE_TZ2), ("tz2_band", _core.TUNE_TZ2_BAND)):
        assert _core.get_tuning(key) == res.knobs[name]
    # nothing fits in one byte
    with pytest.raises(RuntimeError, match="fits"):
        tune(solver, wav, src, rec, [vp], nshots=8, probe_nt=20, repeats=1, max_bytes=1, install=False)
