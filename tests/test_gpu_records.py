"""GPU parity of the record transposes and the L2 misfit (through the C ABI)
against the numpy oracle (oracle/records.py): transposes bit-exact, the
residual bit-exact, the float64 loss within its accumulation bound of the
exactly rounded sum; and at the full 2-D workload's record size
(32 shots x 3000 samples x 500 receivers) the round trip and the loss
against a float64 reduction."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import records as orc
from sweep_amd import loss as swl
from sweep_amd import records

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

SHAPES = [(1, 1, 1, 1), (1, 3, 37, 131), (1, 2, 64, 64), (2, 2, 33, 65), (5, 1, 70, 200), (3, 4, 1, 300),
          (1, 2, 500, 17)]


@pytest.mark.parametrize("nf,B,nrec,nt", SHAPES)
def test_transposes_bit_exact(nf, B, nrec, nt):
    print(f"[records] transpose nf={nf} B={B} nrec={nrec} nt={nt}", flush=True)
    rng = np.random.default_rng(nf + 7 * nt)
    raw = rng.standard_normal((nf, B, nrec, nt) if nf > 1 else (B, nrec, nt)).astype(np.float32)
    can = records.to_canonical(torch.from_numpy(raw).to(DEV))
    assert can.shape == (B, nt, nrec, nf)
    assert np.array_equal(can.cpu().numpy(), orc.to_canonical(raw))
    back = records.to_raw(can, raw.ndim)
    assert np.array_equal(back.cpu().numpy(), raw)
    # to_raw of an independent canonical array
    c2 = rng.standard_normal((B, nt, nrec, nf)).astype(np.float32)
    r2 = records.to_raw(torch.from_numpy(c2).to(DEV), raw.ndim)
    assert np.array_equal(r2.cpu().numpy(), orc.to_raw(c2, raw.ndim))


def test_empty_records():
    z = torch.zeros((2, 0, 10), device=DEV)
    assert records.to_canonical(z).shape == (2, 10, 0, 1)
    assert records.to_raw(torch.zeros((2, 10, 0, 1), device=DEV), 3).shape == (2, 0, 10)


@pytest.mark.parametrize("nf,B,nrec,nt", [(1, 2, 37, 131), (2, 3, 65, 90), (1, 1, 1, 1)])
@pytest.mark.parametrize("masked", [False, True])
def test_misfit_matches_oracle(nf, B, nrec, nt, masked):
    print(f"[records] misfit nf={nf} B={B} nrec={nrec} nt={nt} masked={masked}", flush=True)
    rng = np.random.default_rng(B * nrec + nt)
    pred = rng.standard_normal((B, nt, nrec, nf)).astype(np.float32)
    obs = rng.standard_normal((B, nt, nrec, nf)).astype(np.float32)
    mask = (rng.random((B, nrec)) > 0.3).astype(np.float32) if masked else None
    p = torch.from_numpy(pred).to(DEV).requires_grad_(True)
    o = torch.from_numpy(obs).to(DEV)
    m = torch.from_numpy(mask).to(DEV) if masked else None
    loss = swl.l2_misfit(p, o, m)
    assert loss.dtype == torch.float64
    want = orc.l2_misfit(pred, obs, mask)
    n = pred.size
    assert abs(loss.item() - want) <= 4 * n * 2.0 ** -53 * want + 1e-300
    loss.backward()
    assert np.array_equal(p.grad.cpu().numpy(), orc.residual(pred, obs, mask))
    again = swl.l2_misfit(p.detach(), o, m)
    assert again.item() == loss.item()                     # fixed-order sum
    if masked:
        dead = mask == 0
        assert np.all(p.grad.cpu().numpy()[np.broadcast_to(dead[:, None, :, None], pred.shape)] == 0)


def test_misfit_scales_the_gradient():
    rng = np.random.default_rng(3)
    pred = torch.from_numpy(rng.standard_normal((2, 50, 20, 1)).astype(np.float32)).to(DEV).requires_grad_(True)
    obs = torch.zeros_like(pred)
    (3.0 * swl.l2_misfit(pred, obs)).backward()
    assert torch.equal(pred.grad, pred.detach() * 3.0)


def test_full_size_round_trip_and_loss():
    """The 2-D workload's records: 32 shots, 3000 samples, 500 receivers."""
    print("[records] full size 32 x 3000 x 500", flush=True)
    g = torch.Generator(device=DEV).manual_seed(11)
    raw = torch.randn((32, 500, 3000), device=DEV, generator=g)
    can = records.to_canonical(raw)
    assert torch.equal(records.to_raw(can, 3), raw)
    # spot check against torch's permute
    assert torch.equal(can[..., 0], raw.permute(0, 2, 1))
    obs = torch.randn(can.shape, device=DEV, generator=g)
    mask = (torch.rand((32, 500), device=DEV, generator=g) > 0.2).float()
    loss = swl.l2_misfit(can, obs, mask).item()
    r = ((can - obs) * mask[:, None, :, None]).double()
    want = 0.5 * float((r * r).sum())
    assert loss == pytest.approx(want, rel=1e-11)
