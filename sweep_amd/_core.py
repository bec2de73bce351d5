"""ctypes front end of ``libswp_core.so`` (the C ABI in ``include/sweep_amd.h``).

Plays the part of the reference's ``sweep._C`` / ``backend/c/loader.py``:
dlopen once under a lock, refuse a library whose ABI version or struct layout
differs from this mirror (``loader.py:199-228``), set the thread's stream from
torch's current stream around each call (``backend/c/entries.py:89-102``),
and turn a non-zero return into ``RuntimeError(message)`` (``loader.py:273-277``).

There is no fallback: without the built library every entry raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libswp_core.so"
if os.environ.get("SWP_CORE_LIBRARY"):
    # another build of the same sources (the tests' no-contraction parity build)
    _LIB_PATH = Path(os.environ["SWP_CORE_LIBRARY"]).resolve()
ABI_VERSION = 6
MISFIT_PARTIALS = 1024   # SWP_MISFIT_PARTIALS

_F = ctypes.POINTER(ctypes.c_float)
_I32 = ctypes.POINTER(ctypes.c_int32)

MEM_NONE, MEM_FULL, MEM_BOUNDARY, MEM_CKPT, MEM_BINOMIAL = 0, 1, 2, 3, 4
FLAG_ILLUMINATION, FLAG_BF16_STRIPS, FLAG_GRAPH = 1, 2, 4


class SwpProblem(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32), ("nz", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("B", ctypes.c_int32), ("M", ctypes.c_int32),
        ("nt", ctypes.c_int32), ("abcn", ctypes.c_int32), ("pad", ctypes.c_int32 * 6),
        ("dt", ctypes.c_float), ("dz", ctypes.c_float), ("dy", ctypes.c_float),
        ("dx", ctypes.c_float), ("lap_coes", _F), ("grad_coes", _F), ("pml", _F * 12),
        ("models", _F * 3), ("model_batched", ctypes.c_int32), ("nsrc", ctypes.c_int32),
        ("nrec", ctypes.c_int32), ("source", _F), ("sources_loc", _I32),
        ("receivers_loc", _I32), ("memory", ctypes.c_int32), ("ckpt_interval", ctypes.c_int32),
        ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32), ("nsrc_fields", ctypes.c_int32),
        ("nrec_fields", ctypes.c_int32), ("src_fields", ctypes.c_int32 * 5), ("rec_fields", ctypes.c_int32 * 5),
        ("tape", ctypes.c_void_p),
        ("tape_bytes", ctypes.c_int64), ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_int64),
    ]


class SwpForwardIO(ctypes.Structure):
    _fields_ = [("record", _F)]


class SwpBackwardIO(ctypes.Structure):
    _fields_ = [("adjoint_source", _F), ("grad_source", _F), ("grad_models", _F * 3),
                ("illum_source", _F), ("illum_receiver", _F)]


def _layout_table():
    """sizeof / offsetof in the order swp_layout_probe reports them."""
    out = [ctypes.sizeof(SwpProblem)]
    out += [getattr(SwpProblem, name).offset for name, _ in SwpProblem._fields_]
    out += [ctypes.sizeof(SwpForwardIO), SwpForwardIO.record.offset]
    out += [ctypes.sizeof(SwpBackwardIO)]
    out += [getattr(SwpBackwardIO, name).offset for name, _ in SwpBackwardIO._fields_]
    return out


_lock = threading.RLock()
_lib = None
_entries = None

EXPORTS = (
    "swp_core_abi_version", "swp_layout_count", "swp_layout_probe", "swp_entry_count",
    "swp_entry_name", "swp_entry_kind", "swp_entry_available", "swp_core_device", "swp_query",
    "swp_call", "swp_set_stream", "swp_get_stream", "swp_set_timing", "swp_last_kernel_stats",
    "swp_revolve_schedule", "swp_revolve_spine", "swp_record_to_canonical", "swp_record_to_raw",
    "swp_l2_misfit", "swp_l2_misfit_noalloc",
    "swp_nn_conv_fwd", "swp_nn_conv_wgrad_scratch", "swp_nn_conv_wgrad", "swp_nn_conv_flip", "swp_nn_relu_bwd",
    "swp_nn_pool2_fwd", "swp_nn_pool2_bwd_acc", "swp_nn_up2_fwd", "swp_nn_up2_bwd", "swp_nn_pack_input",
    "swp_nn_mse", "swp_nn_adam", "swp_nn_to_bf16", "swp_set_tuning", "swp_get_tuning",
    "swp_census_enable", "swp_census", "swp_set_graphs", "swp_graph_stats", "swp_graph_clear",
)
TUNE_TZ2, TUNE_TZ2_BAND, TUNE_TZ2_RECON, TUNE_TZ3 = 0, 1, 2, 3


def library_path() -> Path:
    return _LIB_PATH


def _declare(L):
    L.swp_core_abi_version.restype = ctypes.c_int
    L.swp_layout_count.restype = ctypes.c_int64
    L.swp_layout_probe.restype = ctypes.c_int64
    L.swp_layout_probe.argtypes = [ctypes.c_int64]
    L.swp_entry_count.restype = ctypes.c_int
    L.swp_entry_name.restype = ctypes.c_char_p
    L.swp_entry_name.argtypes = [ctypes.c_int]
    L.swp_entry_kind.restype = ctypes.c_int
    L.swp_entry_kind.argtypes = [ctypes.c_int]
    L.swp_entry_available.restype = ctypes.c_int
    L.swp_entry_available.argtypes = [ctypes.c_int]
    L.swp_core_device.restype = ctypes.c_char_p
    L.swp_query.restype = ctypes.c_int
    L.swp_query.argtypes = [ctypes.c_int, ctypes.POINTER(SwpProblem), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int]
    L.swp_call.restype = ctypes.c_int
    L.swp_call.argtypes = [ctypes.c_int, ctypes.POINTER(SwpProblem), ctypes.c_void_p,
                           ctypes.c_char_p, ctypes.c_int]
    L.swp_set_stream.restype = None
    L.swp_set_stream.argtypes = [ctypes.c_void_p]
    L.swp_get_stream.restype = ctypes.c_void_p
    L.swp_set_timing.restype = None
    L.swp_set_timing.argtypes = [ctypes.c_int]
    L.swp_last_kernel_stats.restype = ctypes.c_int
    L.swp_last_kernel_stats.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int64)]
    for f in (L.swp_revolve_schedule, L.swp_revolve_spine):
        f.restype = ctypes.c_int64
        f.argtypes = [ctypes.c_int32, ctypes.c_int32, _I32, ctypes.c_int64]
    _i = ctypes.c_int32
    for f in (L.swp_record_to_canonical, L.swp_record_to_raw):
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _i, _i, _i, _i, ctypes.c_char_p, ctypes.c_int]
    _p, _e, _c = ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int
    _i64, _f, _d = ctypes.c_int64, ctypes.c_float, ctypes.c_double
    nn_sigs = {
        "swp_nn_conv_fwd": [_p, _i, _i, _i, _i, _i, _p, _i, _i, _p, _i, _p, _i, _i],
        "swp_nn_conv_wgrad": [_p, _i, _i, _i, _i, _i, _p, _i, _i, _i, _p, _p, _p],
        "swp_nn_conv_flip": [_p, _i, _i, _i, _p],
        "swp_nn_relu_bwd": [_p, _i, _p, _i, _i64, _i],
        "swp_nn_pool2_fwd": [_p, _i, _i, _i, _i, _i, _p, _i],
        "swp_nn_pool2_bwd_acc": [_p, _i, _i, _i, _i, _i, _p, _i],
        "swp_nn_up2_fwd": [_p, _i, _i, _i, _i, _i, _p, _i],
        "swp_nn_up2_bwd": [_p, _i, _i, _i, _i, _i, _p, _i],
        "swp_nn_pack_input": [_p, _p, _i, _i64, _p, _i],
        "swp_nn_mse": [_p, _i, _p, _p, _i64, _d, _p, _i, _p, _p],
        "swp_nn_adam": [_p, _p, _p, _p, _p, _i64, _f, _f, _f, _f, _i],
        "swp_nn_to_bf16": [_p, _p, _i64],
    }
    for name, args in nn_sigs.items():
        f = getattr(L, name)
        f.restype = _c
        f.argtypes = args + [_e, _c]
    L.swp_set_tuning.restype = _c
    L.swp_set_tuning.argtypes = [_i, _i]
    L.swp_get_tuning.restype = _i
    L.swp_get_tuning.argtypes = [_i]
    L.swp_nn_conv_wgrad_scratch.restype = _i64
    L.swp_nn_conv_wgrad_scratch.argtypes = [_i, _i, _i, _i, _i, _i]
    L.swp_set_graphs.restype = None
    L.swp_set_graphs.argtypes = [_c]
    L.swp_graph_stats.restype = _c
    L.swp_graph_stats.argtypes = [ctypes.POINTER(_i64)] * 4
    L.swp_graph_clear.restype = None
    L.swp_graph_clear.argtypes = []
    L.swp_census_enable.restype = None
    L.swp_census_enable.argtypes = [_c]
    L.swp_census.restype = _i64
    L.swp_census.argtypes = [ctypes.c_char_p, _i64]
    L.swp_l2_misfit_noalloc.restype = ctypes.c_int
    L.swp_l2_misfit_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _i, _i, _i, _i,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
    L.swp_l2_misfit.restype = ctypes.c_int
    L.swp_l2_misfit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, _i, _i, _i, _i,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]


def load():
    """The loaded library, checked; raises when it is missing or drifted."""
    global _lib, _entries
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"libswp_core is not built ({_LIB_PATH}); run `python -m sweep_amd.build` "
                "(or __graft_entry__.build()). There is no fallback implementation.")
        L = ctypes.CDLL(str(_LIB_PATH))
        _declare(L)
        v = L.swp_core_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libswp_core ABI {v} != the Python mirror's {ABI_VERSION}; rebuild")
        want = _layout_table()
        n = L.swp_layout_count()
        got = [L.swp_layout_probe(i) for i in range(n)]
        if got != want:
            raise RuntimeError(f"libswp_core struct layout drifted: core {got} vs mirror {want}; rebuild")
        _entries = {L.swp_entry_name(i).decode(): i for i in range(L.swp_entry_count())}
        _lib = L
        return L


def entries():
    load()
    return dict(_entries)


def entry_id(name: str) -> int:
    e = entries()
    if name not in e:
        raise RuntimeError(f"libswp_core has no entry {name!r}")
    return e[name]


def entry_available(name: str) -> bool:
    L = load()
    return L.swp_entry_available(entry_id(name)) == 1


def _raise(rc, err, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {err.value.decode(errors='replace')}")


def query(name: str, problem: SwpProblem):
    L = load()
    ws = ctypes.c_int64(0)
    tp = ctypes.c_int64(0)
    err = ctypes.create_string_buffer(4096)
    rc = L.swp_query(entry_id(name), ctypes.byref(problem), ctypes.byref(ws), ctypes.byref(tp), err, 4096)
    _raise(rc, err, name)
    return int(ws.value), int(tp.value)


def call(name: str, problem: SwpProblem, io, device=None):
    """Run one entry on torch's current stream of ``device``."""
    import torch
    L = load()
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    err = ctypes.create_string_buffer(4096)
    with torch.cuda.device(dev):
        L.swp_set_stream(ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
        rc = L.swp_call(entry_id(name), ctypes.byref(problem), ctypes.byref(io), err, 4096)
    _raise(rc, err, name)


def set_timing(stride: int):
    """Time one launch in every ``stride`` per kernel class (0: off)."""
    load().swp_set_timing(int(stride))


def kernel_stats(which: int):
    """(total ms, launches) of the last call's kernel class ``which``
    (0 forward step, 1 adjoint step, 2 reconstruction, 3 other)."""
    L = load()
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    if L.swp_last_kernel_stats(which, ctypes.byref(ms), ctypes.byref(n)) != 0:
        raise RuntimeError("swp_last_kernel_stats failed")
    return float(ms.value), int(n.value)


def on_stream(device, fn, *args):
    """fn(*args, err, cap) on torch's current stream of ``device``; raises on
    a non-zero return."""
    import torch
    L = load()
    err = ctypes.create_string_buffer(1024)
    with torch.cuda.device(device):
        L.swp_set_stream(ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream))
        rc = fn(*args, err, 1024)
    _raise(rc, err, fn.__name__)


def set_tuning(key: int, value: int):
    """Launch-shape knob (TUNE_*): value > 0 sets it, 0 restores the default."""
    if load().swp_set_tuning(int(key), int(value)) != 0:
        raise ValueError(f"bad tuning key {key} or value {value}")


def get_tuning(key: int) -> int:
    v = load().swp_get_tuning(int(key))
    if v < 0:
        raise ValueError(f"bad tuning key {key}")
    return int(v)


def set_graphs(mode: int):
    """Time-loop graphs: 1 on, 0 off, -1 the SWP_GRAPHS environment default."""
    load().swp_set_graphs(int(mode))


def graph_stats() -> dict:
    """Captures, replays, eager fallbacks and cached graphs since graph_clear()."""
    v = [ctypes.c_int64(0) for _ in range(4)]
    load().swp_graph_stats(*[ctypes.byref(x) for x in v])
    return dict(zip(("captures", "replays", "eager", "cached"), (int(x.value) for x in v)))


def graph_clear():
    """Free every cached graph and reset the counts."""
    load().swp_graph_clear()


def census_enable(on: bool = True):
    """Clear the launch census and turn it on (or turn it off)."""
    load().swp_census_enable(1 if on else 0)


def census() -> dict:
    """{site: (launches, workgroups)} counted since census_enable(True)."""
    L = load()
    n = L.swp_census(None, 0)
    if n < 0:
        raise RuntimeError("swp_census failed")
    buf = ctypes.create_string_buffer(int(n) + 1)
    L.swp_census(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        site, launches, blocks = line.split()
        out[site] = (int(launches), int(blocks))
    return out


def _pairs(fn, nt: int, slots: int):
    import numpy as np
    n = fn(int(nt), int(slots), None, 0)
    if n < 0:
        raise ValueError(f"bad checkpoint schedule arguments nt={nt}, slots={slots}")
    out = np.zeros((max(n, 1), 2), dtype=np.int32)
    fn(int(nt), int(slots), out.ctypes.data_as(_I32), n)
    return out[:n]


def revolve_schedule(nt: int, slots: int):
    """The binomial checkpoint schedule of Ckpt(mode='recursive', count=slots)
    as an (n, 2) int32 array of (code, arg): 0 step arg, 1 save to slot arg,
    2 load slot arg, 3 re-run and reverse step arg.  Host only (no GPU)."""
    return _pairs(load().swp_revolve_schedule, nt, slots)


def revolve_spine(nt: int, slots: int):
    """(slot, step) snapshots the forward pass of that schedule takes."""
    return _pairs(load().swp_revolve_spine, nt, slots)
