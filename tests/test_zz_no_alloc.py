"""swp_l2_misfit_noalloc (the misfit with its partial sums in the library's
static device array) against swp_l2_misfit (partials in a caller buffer),
through the C ABI: loss and residual bit-identical, on two calls in a row,
the second smaller than the first (so the first call's partial sums are still
in the array when the second runs), with and without a mask.  The loss is
also checked against the exactly rounded sum (math.fsum) of the residual's
squares within 4 n 2^-53 relative.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _records(rng, dev, B, nt, nrec, nf, masked):
    pred = torch.tensor(rng.standard_normal((B, nt, nrec, nf)).astype(np.float32), device=dev)
    obs = torch.tensor(rng.standard_normal((B, nt, nrec, nf)).astype(np.float32), device=dev)
    mask = None
    if masked:
        m = (rng.random((B, nrec)) > 0.3).astype(np.float32)
        mask = torch.tensor(m, device=dev)
    return pred, obs, mask


def _call(fn, args, dev):
    from sweep_amd import _core
    L = _core.load()
    err = ctypes.create_string_buffer(1024)
    with torch.cuda.device(dev):
        L.swp_set_stream(ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        rc = fn(*args, err, 1024)
    assert rc == 0, err.value.decode()


def _both(dev, pred, obs, mask):
    from sweep_amd import _core
    L = _core.load()
    B, nt, nrec, nf = pred.shape
    mp = mask.data_ptr() if mask is not None else None
    res_a = torch.empty_like(pred)
    res_b = torch.empty_like(pred)
    loss_a = torch.zeros((), dtype=torch.float64, device=dev)
    loss_b = torch.zeros((), dtype=torch.float64, device=dev)
    partials = torch.zeros(_core.MISFIT_PARTIALS, dtype=torch.float64, device=dev)
    _call(L.swp_l2_misfit, (pred.data_ptr(), obs.data_ptr(), mp, nf, B, nrec, nt, res_a.data_ptr(),
                            loss_a.data_ptr(), partials.data_ptr()), dev)
    _call(L.swp_l2_misfit_noalloc, (pred.data_ptr(), obs.data_ptr(), mp, nf, B, nrec, nt, res_b.data_ptr(),
                                    loss_b.data_ptr()), dev)
    torch.cuda.synchronize()
    return res_a.cpu().numpy(), float(loss_a.item()), res_b.cpu().numpy(), float(loss_b.item())


@pytest.mark.parametrize("masked", [False, True])
def test_noalloc_misfit_matches_the_workspace_entry(masked):
    dev = _dev()
    rng = np.random.default_rng(11 if masked else 10)
    print(f"[noalloc] masked={masked}: call 1 (4 x 3000 x 200 x 1), call 2 (2 x 500 x 37 x 2)", flush=True)
    for shape in ((4, 3000, 200, 1), (2, 500, 37, 2)):
        pred, obs, mask = _records(rng, dev, *shape, masked)
        res_a, loss_a, res_b, loss_b = _both(dev, pred, obs, mask)
        assert np.array_equal(res_a, res_b), shape
        assert loss_a == loss_b, (shape, loss_a, loss_b)
        exact = 0.5 * math.fsum(float(v) * float(v) for v in res_b.astype(np.float64).ravel())
        n = res_b.size
        assert abs(loss_b - exact) <= 4 * n * 2.0 ** -53 * exact, (shape, loss_b, exact)
        print(f"[noalloc] {shape} loss {loss_b:.17g} (fsum {exact:.17g})", flush=True)


def test_noalloc_misfit_rejects_bad_shapes():
    dev = _dev()
    from sweep_amd import _core
    L = _core.load()
    x = torch.zeros(8, device=dev)
    loss = torch.zeros((), dtype=torch.float64, device=dev)
    err = ctypes.create_string_buffer(1024)
    rc = L.swp_l2_misfit_noalloc(x.data_ptr(), x.data_ptr(), None, 9, 1, 1, 1, None, loss.data_ptr(), err, 1024)
    assert rc != 0 and err.value, "nfield 9 must be refused"
