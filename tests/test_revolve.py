"""The binomial checkpoint schedule of Ckpt(mode='recursive', count=s), host
side (no GPU): libswp_core's swp_revolve_schedule / swp_revolve_spine.

Pinned by what the schedule must do, simulated step by step:
* every REVERSE(t) finds the working state at step t, and t runs nt-1 .. 0;
* every LOAD reads a slot holding a state, every STEP advances by one;
* at most s slots are used (slot 0 holds step 0 from the forward pass);
* no step is re-run more than r times, r the smallest integer with
  C(s + r, s) >= nt (Griewank's repetition bound for s checkpoints);
* with s = 1 the schedule is the quadratic one (nt (nt + 1) / 2 step runs),
  with s >= nt it re-runs each step once (every state is a snapshot).
"""
from __future__ import annotations

from math import comb

import numpy as np
import pytest

from sweep_amd import _core

STEP, SAVE, LOAD, REVERSE = 0, 1, 2, 3


def _simulate(nt, s):
    ops = _core.revolve_schedule(nt, s)
    spine = _core.revolve_spine(nt, s)
    assert spine[0].tolist() == [0, 0]
    slots = {int(k): int(t) for k, t in spine}
    assert max(slots) < s
    cur, expect = None, nt - 1
    runs = np.zeros(nt, dtype=np.int64)
    for code, arg in ops:
        if code == STEP:
            assert cur == arg
            runs[arg] += 1
            cur += 1
        elif code == SAVE:
            assert 0 <= arg < s and cur is not None
            slots[int(arg)] = cur
        elif code == LOAD:
            assert int(arg) in slots
            cur = slots[int(arg)]
        else:
            assert code == REVERSE and cur == arg == expect
            runs[arg] += 1
            expect -= 1
            cur += 1
    assert expect == -1
    return runs


def _r(nt, s):
    r = 0
    while comb(s + r, s) < nt:
        r += 1
    return max(r, 1)


@pytest.mark.parametrize("nt", [1, 2, 5, 17, 100, 400, 3000])
@pytest.mark.parametrize("s", [1, 2, 3, 7, 40, 100])
def test_schedule_is_a_valid_reversal(nt, s):
    runs = _simulate(nt, s)
    assert runs.min() >= 1
    if s > 1:
        assert runs.max() <= _r(nt, s), (runs.max(), _r(nt, s))


def test_one_slot_is_quadratic_and_enough_slots_is_linear():
    assert _simulate(50, 1).sum() == 50 * 51 // 2
    assert _simulate(50, 50).tolist() == [1] * 50
    assert _simulate(50, 80).tolist() == [1] * 50


def test_c2_budget():
    """c2 (nt = 3000) with 100 slots: every step re-run at most twice."""
    runs = _simulate(3000, 100)
    assert runs.max() <= 2 and runs.sum() < 2 * 3000


def test_bad_arguments():
    with pytest.raises(ValueError):
        _core.revolve_schedule(10, 0)
