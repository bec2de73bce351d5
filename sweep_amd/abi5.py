"""A host binding of libswp_core through the reference's C ABI, version 5.

``include/sweep_abi5.h`` exports the ``sweep_*`` functions the reference's
ctypes loader declares (``backend/c/loader.py:52-79``) and takes its
``ForwardInputCore`` / ``BackwardInputCore`` (``csrc/core/input_core.h:10-132``).
This module is the Python side of that boundary for a host that wants it:
ctypes mirrors of the structs (the layout ``sweep_layout_probe`` reports),
:func:`buf_of` (a torch tensor as a ``Buf``), and :func:`load`, which checks
the ABI version and the layout probe the way the reference's loader does
(``loader.py:91-118``) before handing the library out.  The product path,
``sweep_amd.propagator``, calls the ``swp_*`` functions instead; both run the
same kernels.

Buffers differ from the reference in one way (the core never allocates and
keeps its own layout): bind one workspace as ``forward_workspace[0]`` /
``adjoint_workspace[0]`` and one forward->backward tape where the reference
keeps that memory mode's state, the same tensor on both calls --
``u_allt_out`` / ``u_forward`` (full storage), ``boundary_gpu[0]`` (boundary
saving), ``checkpoints[0]`` (checkpointing) -- sized by :func:`query_bytes`.
"""
from __future__ import annotations

import ctypes

from . import _core

ABI_VERSION = 5
BUF_MAX_DIMS = 8
MAX_GRADS = 32
# storage-dtype tags of Buf.dtype_ (BoundaryDtype)
FP32, FP16, BF16, INT8 = 0, 1, 2, 3


class Buf(ctypes.Structure):
    _fields_ = [
        ("data_", ctypes.c_void_p),
        ("sizes_", ctypes.c_int64 * BUF_MAX_DIMS),
        ("strides_", ctypes.c_int64 * BUF_MAX_DIMS),
        ("numel_", ctypes.c_int64),
        ("ndim_", ctypes.c_int32),
        ("elem_size_", ctypes.c_int32),
        ("dtype_", ctypes.c_int32),
        ("defined_", ctypes.c_bool),
        ("is_cuda_", ctypes.c_bool),
        ("device_", ctypes.c_int32),
    ]


class BufList(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.POINTER(Buf)),
        ("n", ctypes.c_int64),
    ]


class IntSpan(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.POINTER(ctypes.c_int32)),
        ("n", ctypes.c_int64),
    ]


class FloatSpan(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.POINTER(ctypes.c_float)),
        ("n", ctypes.c_int64),
    ]


class CStrList(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.POINTER(ctypes.c_char_p)),
        ("n", ctypes.c_int64),
    ]


class GradList(ctypes.Structure):
    _fields_ = [
        ("items", Buf * MAX_GRADS),
        ("n", ctypes.c_int64),
    ]


class ForwardOutputCore(ctypes.Structure):
    _fields_ = [
        ("wavefield", Buf),
        ("last_two", Buf),
        ("record", Buf),
    ]


class RTMOutputCore(ctypes.Structure):
    _fields_ = [
        ("source_illumination", Buf),
        ("receiver_illumination", Buf),
        ("adcig", Buf),
    ]


class BackwardOutputCore(ctypes.Structure):
    _fields_ = [
        ("grads", GradList),
        ("source_illumination", Buf),
        ("receiver_illumination", Buf),
        ("adcig", Buf),
    ]


class ForwardInputCore(ctypes.Structure):
    _fields_ = [
        ("models", BufList),
        ("source", Buf),
        ("lap_coes", Buf),
        ("grad_coes", Buf),
        ("M", ctypes.c_int32),
        ("abcn", ctypes.c_int32),
        ("sources_loc", Buf),
        ("receivers_loc", Buf),
        ("source_field_indices", IntSpan),
        ("receiver_field_indices", IntSpan),
        ("pml_vals", BufList),
        ("wavefields", BufList),
        ("forward_workspace", BufList),
        ("derived_models", BufList),
        ("last_two", Buf),
        ("boundary_cpu", BufList),
        ("boundary_gpu", BufList),
        ("boundary_staging", BufList),
        ("boundary_compress", BufList),
        ("boundary_disk_files", CStrList),
        ("checkpoints", BufList),
        ("checkpoint_steps", Buf),
        ("save_all_wavefields", ctypes.c_bool),
        ("use_boundary_saving", ctypes.c_bool),
        ("use_checkpoint", ctypes.c_bool),
        ("use_recursive_checkpoint", ctypes.c_bool),
        ("checkpoint_on_cpu", ctypes.c_bool),
        ("boundary_on_cpu", ctypes.c_bool),
        ("boundary_on_disk", ctypes.c_bool),
        ("boundary_disk_async_read", ctypes.c_bool),
        ("use_pinned_memory", ctypes.c_bool),
        ("free_surface", ctypes.c_bool),
        ("fs_faces", ctypes.c_int32),
        ("pad_lo", IntSpan),
        ("pad_hi", IntSpan),
        ("eq_aux", BufList),
        ("topo_rows", Buf),
        ("has_topo", ctypes.c_bool),
        ("topo_category", Buf),
        ("use_apm", ctypes.c_bool),
        ("nt", ctypes.c_uint32),
        ("dt", ctypes.c_float),
        ("spacing", FloatSpan),
        ("transfer_interval", ctypes.c_int32),
        ("boundary_ring_buffers", ctypes.c_int32),
        ("boundary_tail_steps", ctypes.c_int32),
        ("boundary_session", ctypes.c_void_p),
        ("checkpoint_interval", ctypes.c_int32),
        ("checkpoint_count", ctypes.c_int32),
        ("it_begin", ctypes.c_int32),
        ("it_end", ctypes.c_int32),
        ("record_out", Buf),
        ("u_allt_out", Buf),
        ("step_phase", ctypes.c_int32),
        ("cut_face_mask", ctypes.c_int32),
    ]


class BackwardInputCore(ctypes.Structure):
    _fields_ = [
        ("u_forward", Buf),
        ("u_boundary", BufList),
        ("u_last_two", Buf),
        ("checkpoints", BufList),
        ("checkpoint_replay", BufList),
        ("checkpoint_steps", Buf),
        ("adjoint_wavefields", BufList),
        ("forward_wavefields", BufList),
        ("adjoint_workspace", BufList),
        ("derived_models", BufList),
        ("boundary_cpu", BufList),
        ("boundary_gpu", BufList),
        ("boundary_staging", BufList),
        ("boundary_compress", BufList),
        ("boundary_disk_files", CStrList),
        ("models", BufList),
        ("adjoint_source", Buf),
        ("forward_source", Buf),
        ("lap_coes", Buf),
        ("grad_coes", Buf),
        ("M", ctypes.c_int32),
        ("abcn", ctypes.c_int32),
        ("adjoint_sources_loc", Buf),
        ("forward_sources_loc", Buf),
        ("source_field_indices", IntSpan),
        ("receiver_field_indices", IntSpan),
        ("pml_vals", BufList),
        ("eq_aux", BufList),
        ("nt", ctypes.c_uint32),
        ("dt", ctypes.c_float),
        ("spacing", FloatSpan),
        ("free_surface", ctypes.c_bool),
        ("fs_faces", ctypes.c_int32),
        ("pad_lo", IntSpan),
        ("pad_hi", IntSpan),
        ("topo_rows", Buf),
        ("has_topo", ctypes.c_bool),
        ("topo_category", Buf),
        ("use_apm", ctypes.c_bool),
        ("checkpoint_on_cpu", ctypes.c_bool),
        ("boundary_on_cpu", ctypes.c_bool),
        ("boundary_on_disk", ctypes.c_bool),
        ("boundary_disk_async_read", ctypes.c_bool),
        ("use_pinned_memory", ctypes.c_bool),
        ("transfer_interval", ctypes.c_int32),
        ("boundary_ring_buffers", ctypes.c_int32),
        ("boundary_tail_steps", ctypes.c_int32),
        ("boundary_session", ctypes.c_void_p),
        ("checkpoint_interval", ctypes.c_int32),
        ("checkpoint_count", ctypes.c_int32),
        ("compute_illumination", ctypes.c_bool),
        ("bw_it_begin", ctypes.c_int32),
        ("bw_it_end", ctypes.c_int32),
        ("step_phase", ctypes.c_int32),
        ("grads_out", BufList),
        ("illum_out", BufList),
        ("adcig_out", Buf),
        ("grad_split_iii_out", Buf),
        ("cut_face_mask", ctypes.c_int32),
        ("compute_adcig", ctypes.c_bool),
        ("adcig_max_lag", ctypes.c_int32),
    ]


LAYOUT = [Buf, BufList, IntSpan, FloatSpan, CStrList, GradList, ForwardOutputCore, RTMOutputCore,
          BackwardOutputCore, ForwardInputCore, BackwardInputCore]


def layout_table():
    """sizeof / offsetof of this mirror, struct by struct in the order the
    reference's probe lists them (its first entries)."""
    out = []
    for S in LAYOUT:
        out.append(ctypes.sizeof(S))
        out += [getattr(S, name).offset for name, _ in S._fields_]
    return out


_lib = None


def load():
    """libswp_core with the sweep_* signatures declared, after the ABI guard:
    version 5 and the mirror's sizes/offsets equal to the probe's entries."""
    global _lib
    if _lib is not None:
        return _lib
    L = _core.load()
    i64, i32, vp, cp = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p
    sigs = {
        "sweep_core_abi_version": ([], i32), "sweep_layout_count": ([], i64), "sweep_layout_probe": ([i64], i64),
        "sweep_entry_count": ([], i32), "sweep_entry_name": ([i32], cp), "sweep_entry_kind": ([i32], i32),
        "sweep_entry_available": ([i32], i32), "sweep_core_device": ([], cp),
        "sweep_call": ([i32, vp, vp, cp, i32], i32), "sweep_set_stream": ([vp], None), "sweep_get_stream": ([], vp),
        "sweep_query_bytes": ([i32, vp, ctypes.POINTER(i64), ctypes.POINTER(i64), cp, i32], i32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes, f.restype = args, res
    if L.sweep_core_abi_version() != ABI_VERSION:
        raise RuntimeError(f"core ABI {L.sweep_core_abi_version()} != {ABI_VERSION}")
    mine = layout_table()
    got = [L.sweep_layout_probe(i) for i in range(len(mine))]
    if got != mine:
        raise RuntimeError("ForwardInputCore/BackwardInputCore mirror disagrees with the core's layout probe")
    _lib = L
    return L


def buf_of(t, dtype=FP32):
    """A Buf describing tensor ``t`` (None: an unbound Buf)."""
    b = Buf()
    if t is None:
        return b
    b.data_ = t.data_ptr()
    b.ndim_ = t.dim()
    for k in range(t.dim()):
        b.sizes_[k] = t.shape[k]
        b.strides_[k] = t.stride(k)
    b.numel_ = t.numel()
    b.elem_size_ = t.element_size()
    b.dtype_ = dtype
    b.defined_ = True
    b.is_cuda_ = t.is_cuda
    b.device_ = t.device.index if t.is_cuda else -1
    return b


def buf_list(tensors):
    """(BufList, keep-alive array) of a list of tensors."""
    arr = (Buf * max(len(tensors), 1))(*[buf_of(t) for t in tensors])
    return BufList(ctypes.cast(arr, ctypes.POINTER(Buf)), len(tensors)), arr


def int_span(values):
    arr = (ctypes.c_int32 * max(len(values), 1))(*values)
    return IntSpan(ctypes.cast(arr, ctypes.POINTER(ctypes.c_int32)), len(values)), arr


def float_span(values):
    arr = (ctypes.c_float * max(len(values), 1))(*values)
    return FloatSpan(ctypes.cast(arr, ctypes.POINTER(ctypes.c_float)), len(values)), arr


def entry_id(name):
    L = load()
    for i in range(L.sweep_entry_count()):
        if L.sweep_entry_name(i).decode() == name:
            return i
    raise RuntimeError(f"no entry {name!r}")


def query_bytes(name, params):
    """(workspace bytes, tape bytes) the call ``params`` describes needs."""
    L = load()
    ws, tp = ctypes.c_int64(0), ctypes.c_int64(0)
    err = ctypes.create_string_buffer(4096)
    if L.sweep_query_bytes(entry_id(name), ctypes.addressof(params), ctypes.byref(ws), ctypes.byref(tp), err,
                           4096) != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
    return int(ws.value), int(tp.value)


def call(name, params, stream=0):
    """sweep_call(entry, &params, &out) on ``stream`` (a hipStream_t handle)."""
    L = load()
    out = BackwardOutputCore() if L.sweep_entry_kind(entry_id(name)) == 1 else ForwardOutputCore()
    err = ctypes.create_string_buffer(4096)
    prev = L.sweep_get_stream()
    L.sweep_set_stream(ctypes.c_void_p(stream))
    try:
        rc = L.sweep_call(entry_id(name), ctypes.addressof(params), ctypes.addressof(out), err, 4096)
    finally:
        L.sweep_set_stream(ctypes.c_void_p(prev))
    if rc != 0:
        raise RuntimeError(err.value.decode(errors="replace"))
