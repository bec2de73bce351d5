"""The FWI line search, host logic only (no GPU)."""
from __future__ import annotations

import pytest

from sweep_amd.fwi import line_search


def test_expands_while_the_misfit_keeps_falling():
    calls = []

    def f(s):
        calls.append(s)
        return (s - 3.0) ** 2

    r = line_search(f, f0=9.0, step0=1.0, max_doublings=3)
    assert calls == [1.0, 2.0, 4.0]          # 4 is no better than 2: stop
    assert r.step == 2.0 and r.misfit == 1.0 and r.evals == 3


def test_halves_until_the_misfit_falls():
    r = line_search(lambda s: (s - 0.3) ** 2, f0=0.09, step0=4.0, max_halvings=5)
    assert r.step == 0.5 and r.misfit < 0.09
    assert [s for s, _ in r.trials] == [4.0, 2.0, 1.0, 0.5]


def test_no_step_when_nothing_lowers_the_misfit():
    r = line_search(lambda s: 1.0 + s, f0=1.0, step0=1.0, max_halvings=3)
    assert r.step == 0.0 and r.misfit == 1.0 and r.evals == 4


def test_bad_first_step():
    with pytest.raises(ValueError):
        line_search(lambda s: s, 0.0, 0.0)
