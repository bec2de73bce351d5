"""The reference's C ABI (version 5) over libswp_core, on the CPU: every
symbol the reference's loader declares is exported, the ABI version and the
layout probe are the reference's, the library's struct mirror agrees with
the probe, and -- where the reference tree is present -- with the
reference's own input_core.h compiled by g++.  No kernel runs here."""
from __future__ import annotations

import ctypes
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from sweep_amd import _core, abi5

ROOT = Path(__file__).resolve().parents[1]
TABLE = ROOT / "sweep_amd" / "csrc" / "sweep_abi5_layout.inc"
HEADER = ROOT / "include" / "sweep_abi5.h"
REF_CORE = Path("/root/reference/src/sweep/csrc/core")

# The functions the reference's loader declares before anything else
# (backend/c/loader.py:52-79); a core missing one is refused by name.
LOADER_SYMBOLS = [
    "sweep_core_abi_version", "sweep_core_device", "sweep_entry_count", "sweep_entry_name", "sweep_entry_kind",
    "sweep_entry_available", "sweep_call", "sweep_sem_call", "sweep_entry_struct",
    "sweep_forward_runner_create", "sweep_forward_runner_run", "sweep_forward_runner_device_index",
    "sweep_forward_runner_destroy", "sweep_backward_runner_create", "sweep_backward_runner_run",
    "sweep_backward_runner_device_index", "sweep_backward_runner_destroy", "sweep_session_create",
    "sweep_session_finish", "sweep_session_used", "sweep_session_destroy", "sweep_set_stream", "sweep_get_stream",
    "sweep_visco_fft_workspace_bytes", "sweep_layout_count", "sweep_layout_probe", "sweep_nvcomp_load",
    "sweep_nvcomp_loaded", "sweep_nvcomp_piece_bytes", "sweep_nvcomp_max_piece_out",
    "sweep_nvcomp_decompress_temp_bytes",
]


def _table():
    rows = []
    for line in TABLE.read_text().splitlines()[1:]:
        value, _, name = line.partition(",")
        rows.append((int(value), name.split("//")[1].strip()))
    return rows


def test_exports_every_symbol_the_loader_declares():
    L = ctypes.CDLL(str(_core.library_path()))
    missing = [s for s in LOADER_SYMBOLS if not hasattr(L, s)]
    assert not missing, missing
    declared = set(re.findall(r"SWP_CAPI\s+[\w\s\*]+?\b(sweep_\w+)\s*\(", HEADER.read_text()))
    assert set(LOADER_SYMBOLS) <= declared
    for name in declared:
        assert hasattr(L, name), name


def test_version_probe_and_entries():
    L = abi5.load()
    assert L.sweep_core_abi_version() == 5
    rows = _table()
    assert len(rows) == 2333 and L.sweep_layout_count() == 2333
    assert [L.sweep_layout_probe(i) for i in range(len(rows))] == [v for v, _ in rows]
    assert L.sweep_layout_probe(2333) == -1 and L.sweep_layout_probe(-1) == -1
    # the BoundaryDtype enumerators close the table
    assert rows[-4:] == [(0, "BoundaryDtype::FP32"), (1, "BoundaryDtype::FP16"), (2, "BoundaryDtype::BF16"),
                         (3, "BoundaryDtype::INT8")]
    # the structs sweep_call reads lead the probe and match the Python mirror
    mine = abi5.layout_table()
    assert mine == [v for v, _ in rows[:len(mine)]]
    names = {L.sweep_entry_name(i).decode() for i in range(L.sweep_entry_count())}
    assert {"acoustic2d_forward", "acoustic3d_backward_bs", "elastic2d_backward_recursive_ckpt"} <= names
    assert L.sweep_core_device() == b"cuda"


def test_refusals_come_back_as_messages():
    L = abi5.load()
    err = ctypes.create_string_buffer(256)
    L.sweep_sem_call.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int]
    assert L.sweep_sem_call(0, None, 0, err, 256) != 0 and b"SEM" in err.value
    fin = abi5.ForwardInputCore()
    assert L.sweep_call(0, None, None, err, 256) != 0 and b"null input" in err.value
    assert L.sweep_call(999, ctypes.addressof(fin), None, err, 256) != 0 and b"out of range" in err.value
    assert L.sweep_call(0, ctypes.addressof(fin), None, err, 256) != 0 and b"models" in err.value


@pytest.mark.skipif(not REF_CORE.exists() or shutil.which("g++") is None,
                    reason="needs the reference tree (csrc/core/input_core.h) and g++")
def test_mirror_matches_the_reference_headers(tmp_path):
    """Compile the reference's own input_core.h / outputs.h and print sizeof
    and offsetof of every field the probe lists for the structs sweep_call
    reads; they must equal this core's probe."""
    rows = _table()[:len(abi5.layout_table())]
    lines = []
    for _, name in rows:
        s, _, f = name.partition("::")
        lines.append(f'    std::printf("%lld\\n", (long long)' + (f"offsetof({s}, {f})" if f else f"sizeof({s})") + ");")
    src = tmp_path / "probe.cpp"
    src.write_text("#include <cstddef>\n#include <cstdio>\n#include \"input_core.h\"\n#include \"outputs.h\"\n"
                   "int main() {\n" + "\n".join(lines) + "\n    return 0;\n}\n")
    exe = tmp_path / "probe"
    r = subprocess.run(["g++", "-std=c++17", "-Wno-invalid-offsetof", f"-I{REF_CORE}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [v for v, _ in rows]
