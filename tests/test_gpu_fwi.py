"""Plain illumination-preconditioned FWI on the GPU kernels: a small 2-D
inversion (crosswell-free surface acquisition, 6 shots in 2 batches) whose
misfit must fall at every accepted step, with the preconditioned direction a
descent direction of the gradient."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from sweep_amd import synthetic
from sweep_amd.equations import Acoustic
from sweep_amd.fwi import PlainFWI
from sweep_amd.options import BoundarySaving
from sweep_amd.propagator import PropTorch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _problem():
    shape, nt, dt = (64, 128), 700, 1e-3
    vp_true = synthetic.layered_2d(shape, seed=3, vmin=1500.0, vmax=3000.0, nlayers=4)
    vp0 = synthetic.smoothed(vp_true, sigma=6.0)
    src, rec, mask = synthetic.surface_acquisition(shape, 6, seed=3, nrec_range=(60, 120))
    wav = synthetic.ricker(np.arange(nt) * dt - 0.1, 12.0)
    solver = PropTorch(Acoustic(spatial_order=8, device=DEV), shape=shape, dh=10.0, dt=dt, device=DEV,
                       abcn=20, memory=BoundarySaving())
    batches = [(0, 3), (3, 6)]
    vt = torch.tensor(vp_true, device=DEV)
    with torch.no_grad():
        obs = [solver(wav, src[a:b], rec[a:b], models=[vt]) for a, b in batches]
    masks = [torch.tensor(mask[a:b], device=DEV) for a, b in batches]
    return solver, wav, src, rec, obs, batches, masks, vt, torch.tensor(vp0, device=DEV)


def test_plain_fwi_lowers_the_misfit():
    print("[fwi] plain FWI 64x128, 6 shots, 3 iterations", flush=True)
    solver, wav, src, rec, obs, batches, masks, vt, vp = _problem()
    fwi = PlainFWI(solver, wav, src, rec, obs, batches, masks, vmin=1400.0, vmax=3200.0, first_step=40.0)
    J0, g, S = fwi.gradient(vp)
    assert g.shape == vp.shape and S.shape == vp.shape
    assert torch.isfinite(g).all() and bool((S >= 0).all()) and bool(S.max() > 0)
    assert J0 == pytest.approx(fwi.misfit(vp), rel=1e-6)
    d = fwi.direction(g, S)
    assert float((g.double() * d.double()).sum()) < 0          # a descent direction
    assert float(d.abs().max()) == pytest.approx(1.0)
    err0 = float((vp - vt).norm())
    out = fwi.run(vp, 3)
    h = fwi.history
    assert len(h) == 3
    assert h[0].misfit == pytest.approx(J0, rel=1e-6)
    for r in h:
        print(f"[fwi] J {r.misfit:.6e} -> {r.misfit_after:.6e} step {r.step:.2f} evals {r.evals} "
              f"{r.seconds:.2f} s", flush=True)
        if r.step > 0:
            assert r.misfit_after < r.misfit
    assert h[-1].misfit_after < 0.8 * J0
    assert float(out.min()) >= 1400.0 and float(out.max()) <= 3200.0
    assert float((out - vt).norm()) < err0
