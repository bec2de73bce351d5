"""Record wire format and L2 misfit, host side (no GPU): the oracle's
restatement pinned index by index, and the C ABI's argument checks.

The canonical layout is (B, nt, nrec, nfield) and the raw one
(nfield, B, nrec, nt) (propagator/_c.py:95-186): canonical[b, t, r, f] must
equal raw[f, b, r, t] for every index, which the loops below check on small
ragged shapes, plus the round trip.  The misfit's float64 sum is pinned by
closed forms (a constant residual c over n entries gives 0.5 n c^2 exactly,
a zero mask gives 0, masking receivers equals dropping them)."""
from __future__ import annotations

import ctypes
import itertools

import numpy as np
import pytest

from oracle import records as orc
from sweep_amd import _core


@pytest.mark.parametrize("nf,B,nrec,nt", [(1, 1, 1, 1), (1, 2, 3, 5), (2, 2, 4, 3), (5, 1, 3, 4)])
def test_oracle_layouts_index_by_index(nf, B, nrec, nt):
    rng = np.random.default_rng(nf * 100 + nt)
    raw = rng.standard_normal((nf, B, nrec, nt)).astype(np.float32)
    src = raw[0] if nf == 1 else raw
    can = orc.to_canonical(src)
    assert can.shape == (B, nt, nrec, nf)
    for f, b, r, t in itertools.product(range(nf), range(B), range(nrec), range(nt)):
        assert can[b, t, r, f] == raw[f, b, r, t]
    back = orc.to_raw(can, 3 if nf == 1 else 4)
    assert np.array_equal(back, src)


def test_oracle_misfit_closed_forms():
    B, nt, nrec, nf = 3, 7, 5, 2
    obs = np.zeros((B, nt, nrec, nf), np.float32)
    pred = np.full_like(obs, 0.375)      # exact in binary
    n = obs.size
    assert orc.l2_misfit(pred, obs) == 0.5 * n * 0.375 ** 2
    assert orc.l2_misfit(pred, obs, np.zeros((B, nrec), np.float32)) == 0.0
    rng = np.random.default_rng(5)
    pred = rng.standard_normal(obs.shape).astype(np.float32)
    obs = rng.standard_normal(obs.shape).astype(np.float32)
    keep = np.array([[1, 0, 1, 1, 0], [0, 0, 0, 0, 1], [1, 1, 1, 1, 1]], np.float32)
    masked = orc.l2_misfit(pred, obs, keep)
    dropped = sum(orc.l2_misfit(pred[b:b + 1, :, keep[b] > 0], obs[b:b + 1, :, keep[b] > 0])
                  for b in range(B) if keep[b].any())
    assert masked == pytest.approx(dropped, rel=1e-15)
    r = orc.residual(pred, obs, keep)
    assert np.all(r[0, :, 1] == 0) and np.all(r[1, :, :4] == 0)


def test_abi_rejects_bad_record_shapes():
    L = _core.load()
    err = ctypes.create_string_buffer(256)
    assert L.swp_record_to_canonical(None, None, 6, 1, 1, 1, err, 256) == 1
    assert b"nfield" in err.value
    assert L.swp_record_to_raw(None, None, 0, 1, 1, 1, err, 256) == 1
    assert L.swp_record_to_canonical(None, None, 1, 1, -1, 1, err, 256) == 1
    # an empty record is a no-op (nothing is launched)
    assert L.swp_record_to_canonical(None, None, 1, 4, 0, 10, err, 256) == 0
    assert L.swp_l2_misfit(None, None, None, 1, 1, 1, 1, None, None, None, err, 256) == 1
    assert b"loss" in err.value
