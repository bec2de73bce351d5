"""ORACLE -- test infrastructure only.

ctypes front end of ``oracle/_ref/libref_acoustic.so``: the reference's own
compiled acoustic path (its CPU host core, built from the reference's sources
by ``oracle/build_ref.sh``) driven through ``oracle/ref_harness.cpp`` in
full-storage mode.  Present only where the reference tree was there to build
from; tests that use it skip otherwise.  It pins the C oracle's arithmetic
against the reference's compiled arithmetic.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libref_acoustic.so"

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def available():
    return LIB.exists()


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(str(LIB))
        L.ref_acoustic_run.restype = ctypes.c_int
        L.ref_elastic_run.restype = ctypes.c_int
        L.ref_elastic_run_bs.restype = ctypes.c_int
        L.ref_acoustic2d_run_bs.restype = ctypes.c_int
        L.ref_acoustic3d_run_bs.restype = ctypes.c_int
        L.ref_acoustic_run_illum.restype = ctypes.c_int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(_F) if a is not None else None


def run_illum(P, vp_pad, wav, sloc, rloc, res, bs=False):
    """The reference's backward with compute_illumination: ``dict(rec,
    grad_vp, grad_wavelet, illum_s, illum_r)`` on the runtime grid."""
    ndim = 3 if hasattr(P, "ny") else 2
    grid = np.array([P.nz, P.ny, P.nx] if ndim == 3 else [P.nz, P.nx], dtype=np.int32)
    sp = np.array([P.dh_zyx[k] for k in range(ndim)], dtype=np.float32)
    B, nt = P.B, P.nt
    nsrc, nrec = sloc.shape[1], rloc.shape[1]
    prof = [_f32(p) for p in P.prof]
    parr = (_F * len(prof))(*[_fp(p) for p in prof])
    vp_pad, wav, res = _f32(vp_pad), _f32(wav), _f32(res)
    sloc, rloc = np.ascontiguousarray(sloc, np.int32), np.ascontiguousarray(rloc, np.int32)
    rec = np.zeros((B, nrec, nt), np.float32)
    gv = np.zeros(vp_pad.shape, np.float32)
    gw = np.zeros((B, nsrc, nt), np.float32)
    S = np.zeros(vp_pad.shape, np.float32)
    R = np.zeros(vp_pad.shape, np.float32)
    err = ctypes.create_string_buffer(1024)
    rc = lib().ref_acoustic_run_illum(ndim, 1 if bs else 0, grid.ctypes.data_as(_I), B, P.M, nt, P.abcn,
                                      ctypes.c_float(P.dt), _fp(sp), parr, _fp(vp_pad), _fp(P.lap), _fp(P.grad),
                                      _fp(wav), sloc.ctypes.data_as(_I), nsrc, rloc.ctypes.data_as(_I), nrec,
                                      _fp(res), _fp(rec), _fp(gv), _fp(gw), _fp(S), _fp(R), err, 1024)
    if rc != 0:
        raise RuntimeError("reference core: " + err.value.decode(errors="replace"))
    return dict(rec=rec, grad_vp=gv, grad_wavelet=gw, illum_s=S, illum_r=R)


def run(P, vp_pad, wav, sloc, rloc, res=None, bs=False):
    """Forward (and, with ``res``, the full-storage backward -- the
    boundary-saving one with ``bs``) of the reference for problem
    ``P`` (an ``A2DProblem`` / ``A3DProblem``).  Returns
    ``dict(rec, grad_vp, grad_wavelet)`` on the runtime grid."""
    ndim = 3 if hasattr(P, "ny") else 2
    grid = np.array([P.nz, P.ny, P.nx] if ndim == 3 else [P.nz, P.nx], dtype=np.int32)
    sp = np.array([P.dh_zyx[k] for k in range(ndim)], dtype=np.float32)
    B, nt = P.B, P.nt
    nsrc, nrec = sloc.shape[1], rloc.shape[1]
    prof = [_f32(p) for p in P.prof]
    parr = (_F * len(prof))(*[_fp(p) for p in prof])
    vp_pad = _f32(vp_pad)
    wav, res = _f32(wav), (_f32(res) if res is not None else None)
    sloc, rloc = np.ascontiguousarray(sloc, np.int32), np.ascontiguousarray(rloc, np.int32)
    rec = np.zeros((B, nrec, nt), np.float32)
    gv = np.zeros(vp_pad.shape, np.float32)
    gw = np.zeros((B, nsrc, nt), np.float32)
    err = ctypes.create_string_buffer(1024)
    if bs:
        fn = lib().ref_acoustic2d_run_bs if ndim == 2 else lib().ref_acoustic3d_run_bs
        rc = fn(grid.ctypes.data_as(_I), B, P.M, nt, P.abcn, ctypes.c_float(P.dt),
                                         _fp(sp), parr, _fp(vp_pad), _fp(P.lap), _fp(P.grad), _fp(wav),
                                         sloc.ctypes.data_as(_I), nsrc, rloc.ctypes.data_as(_I), nrec, _fp(res),
                                         _fp(rec), _fp(gv), _fp(gw), err, 1024)
    else:
        rc = lib().ref_acoustic_run(ndim, grid.ctypes.data_as(_I), B, P.M, nt, P.abcn, ctypes.c_float(P.dt),
                                    _fp(sp), parr, _fp(vp_pad), _fp(P.lap), _fp(P.grad), _fp(wav),
                                    sloc.ctypes.data_as(_I), nsrc, rloc.ctypes.data_as(_I), nrec, _fp(res),
                                    _fp(rec), _fp(gv), _fp(gw), err, 1024)
    if rc != 0:
        raise RuntimeError("reference core: " + err.value.decode(errors="replace"))
    return dict(rec=rec, grad_vp=gv, grad_wavelet=gw)


def run_elastic(shape_rt, M, nt, abcn, dt, dh_zx, prof, vp, vs, rho, grad, src, sloc, rloc, res=None, bs=False):
    """Forward (records vx, vz) and, with ``res`` (2, B, nrec, nt), the
    full-storage backward of the reference's elastic 2-D core.  Arrays on
    the runtime grid (B, nz, nx); ``prof`` = [az, bz, azh, bzh, ax, bx, axh,
    bxh]; ``grad`` the (M+1) staggered table; one vz source per shot.
    Returns ``dict(rec (2, B, nrec, nt), grad_vp, grad_vs, grad_rho)``."""
    B = vp.shape[0]
    grid = np.array(shape_rt, dtype=np.int32)
    nsrc, nrec = sloc.shape[1], rloc.shape[1]
    pr = [_f32(p) for p in prof]
    parr = (_F * 8)(*[_fp(p) for p in pr])
    vp, vs, rho, grad, src = (_f32(a) for a in (vp, vs, rho, grad, src))
    res = _f32(res) if res is not None else None
    sloc, rloc = np.ascontiguousarray(sloc, np.int32), np.ascontiguousarray(rloc, np.int32)
    rec = np.zeros((2, B, nrec, nt), np.float32)
    g = [np.zeros(vp.shape, np.float32) for _ in range(3)]
    sp = np.array(dh_zx, dtype=np.float32)
    err = ctypes.create_string_buffer(1024)
    fn = lib().ref_elastic_run_bs if bs else lib().ref_elastic_run
    rc = fn(grid.ctypes.data_as(_I), B, M, nt, abcn, ctypes.c_float(dt), _fp(sp), parr,
                               _fp(vp), _fp(vs), _fp(rho), _fp(grad), _fp(src), sloc.ctypes.data_as(_I), nsrc,
                               rloc.ctypes.data_as(_I), nrec, _fp(res), _fp(rec), _fp(g[0]), _fp(g[1]), _fp(g[2]),
                               err, 1024)
    if rc != 0:
        raise RuntimeError("reference core: " + err.value.decode(errors="replace"))
    return dict(rec=rec, grad_vp=g[0], grad_vs=g[1], grad_rho=g[2])
