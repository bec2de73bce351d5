"""The C-ABI library (CPU side: load, exports, layout guard, entry table,
error path).  No kernel runs here."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from sweep_amd import _core

HEADER = Path(__file__).resolve().parents[1] / "include" / "sweep_amd.h"


def test_library_is_built_and_loads():
    assert _core.library_path().exists(), "run __graft_entry__.build() first"
    _core.load()


def test_exports_every_declared_symbol():
    decl = set(re.findall(r"SWP_CAPI\s+[\w\s\*]+?\b(swp_\w+)\s*\(", HEADER.read_text()))
    assert decl, "no declarations parsed"
    assert decl == set(_core.EXPORTS)
    L = ctypes.CDLL(str(_core.library_path()))
    for name in decl:
        assert hasattr(L, name), name


def test_abi_version_and_layout_probe():
    L = _core.load()
    assert L.swp_core_abi_version() == _core.ABI_VERSION
    n = L.swp_layout_count()
    assert [L.swp_layout_probe(i) for i in range(n)] == _core._layout_table()
    assert L.swp_layout_probe(n) == -1 and L.swp_layout_probe(-1) == -1


def test_entry_table_names_the_plugin_slot():
    names = _core.entries()
    for eq in ("acoustic2d", "acoustic3d", "elastic2d"):
        for op in ("forward", "backward", "backward_bs", "backward_ckpt", "backward_recursive_ckpt"):
            assert f"{eq}_{op}" in names
    L = _core.load()
    assert L.swp_entry_name(len(names)) is None
    assert L.swp_entry_kind(-1) == -1
    assert _core.entry_available("acoustic2d_forward")
    assert L.swp_core_device() == b"cuda"


def test_errors_come_back_as_messages():
    P = _core.SwpProblem()
    P.ndim, P.B, P.nt = 2, 0, 10
    with pytest.raises(RuntimeError, match="need B >= 1"):
        _core.query("acoustic2d_forward", P)
    with pytest.raises(RuntimeError, match="no entry"):
        _core.entry_id("acoustic2d_nonsense")


def test_query_sizes_scale_with_the_problem():
    import numpy as np
    P = _core.SwpProblem()
    P.ndim, P.nz, P.ny, P.nx, P.B, P.M, P.nt, P.abcn = 2, 120, 1, 140, 2, 4, 50, 10
    for k, v in enumerate([10, 10, 0, 0, 10, 10]):
        P.pad[k] = v
    P.dt, P.dz, P.dx = 1e-3, 10.0, 10.0
    lap = np.zeros(5, np.float32)
    P.lap_coes = lap.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    P.grad_coes = lap.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    P.memory = _core.MEM_FULL
    ws, tp = _core.query("acoustic2d_forward", P)
    n = 2 * 120 * 140 * 4
    assert ws >= 9 * n and tp >= 50 * n
    P.memory = _core.MEM_BOUNDARY
    ws2, tp2 = _core.query("acoustic2d_backward_bs", P)
    assert ws2 >= 20 * n and tp2 < tp
    with pytest.raises(RuntimeError, match="memory mode"):
        _core.query("acoustic2d_backward", P)


def test_tuning_table():
    for key in (_core.TUNE_TZ2, _core.TUNE_TZ2_BAND, _core.TUNE_TZ2_RECON, _core.TUNE_TZ3):
        assert _core.get_tuning(key) == 0
        _core.set_tuning(key, 48)
        assert _core.get_tuning(key) == 48
        _core.set_tuning(key, 0)
        assert _core.get_tuning(key) == 0
    with pytest.raises(ValueError):
        _core.set_tuning(7, 1)
    with pytest.raises(ValueError):
        _core.set_tuning(_core.TUNE_TZ3, -2)
    with pytest.raises(ValueError):
        _core.get_tuning(-1)
