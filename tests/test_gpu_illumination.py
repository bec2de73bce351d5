"""GPU source / receiver illumination against the reference core's
compute_illumination (fixtures golden/refcore{2,3}d_illum[_bs]_*.npz from
tests/golden/gen_refcore.py), full storage and boundary saving: the
adjoint source is the fixture's residual; S and R are summed over shots
and cropped to the physical box as PropTorch.collect_illumination returns
them.  Bar 1e-4 (FMA contraction of u_tt and lambda on the GPU)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import GOLDEN, rel_l2

pytestmark = pytest.mark.gpu

CASES = ["refcore2d_illum_o8", "refcore2d_illum_bs_o8", "refcore3d_illum_o4", "refcore3d_illum_bs_o4"]


@pytest.mark.parametrize("case", CASES)
def test_illumination_matches_reference_core(case):
    dev = torch.device("cuda:0")
    print(f"[illum] {case}", flush=True)
    from oracle.oracle import A2DProblem, A3DProblem
    from sweep_amd.equations import Acoustic, Acoustic3D
    from sweep_amd.options import BoundarySaving, Full
    from sweep_amd.propagator import PropTorch
    z = np.load(GOLDEN / f"{case}.npz")
    ndim = len(z["shape"])
    shape = tuple(int(v) for v in z["shape"])
    B, nt, order = int(z["B"]), int(z["nt"]), int(z["order"])
    P = (A2DProblem if ndim == 2 else A3DProblem)(order, shape, float(z["dh"]), float(z["dt"]), nt,
                                                  abcn=int(z["abcn"]), B=B)
    eq = (Acoustic if ndim == 2 else Acoustic3D)(spatial_order=order, device=dev)
    solver = PropTorch(eq, shape=shape, dh=float(z["dh"]), dt=float(z["dt"]), device=dev, abcn=int(z["abcn"]),
                       memory=BoundarySaving() if "_bs_" in case else Full(), compute_illumination=True)
    m = torch.tensor(z["vp"], device=dev, requires_grad=True)
    w = np.broadcast_to(z["wavelet"], (B, nt)).copy()
    pred = solver(w, z["sources"][:, 0, :], z["receivers"], models=[m])
    res = np.ascontiguousarray(np.asarray(z["residual"]).transpose(0, 2, 1)[..., None])
    (pred * torch.tensor(res, device=dev)).sum().backward()
    torch.cuda.synchronize()
    S, R = solver.collect_illumination()
    eS = rel_l2(S.cpu().numpy(), P.crop(z["illum_s"]).sum(0))
    eR = rel_l2(R.cpu().numpy(), P.crop(z["illum_r"]).sum(0))
    eg = rel_l2(m.grad.cpu().numpy(), P.crop(z["grad_vp"]).sum(0))
    print(f"[illum] {case} S {eS:.2e} R {eR:.2e} grad {eg:.2e}", flush=True)
    assert eS <= 1e-4 and eR <= 1e-4
    assert eg <= 1e-4
