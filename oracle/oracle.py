"""ORACLE -- test infrastructure only.

ctypes front end of the C restatement in ``oracle/acoustic2d.c``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, as the checker; the product package never does.

:func:`a2d_setup` restates the reference's host-side problem assembly for the
2-D acoustic path, with numpy:

* runtime grid = physical + PML pad + 2M (``propagator/base.py:323-345``);
* coordinates shifted by the low-side pad + M per axis, (x, z) order
  (``core/geometry.py:48-56``, ``propagator/_c.py:2184-2207``);
* model replicate-padded by pad + M (``utils/torch.py:8-17``);
* CPML profiles on the runtime grid with ``fd_pad = M`` (``_c.py:2199-2200``,
  ``propagator/base.py:1006-1069``), built by ``sweep_amd.pml``, which the
  tests pin bit-exactly against the reference's own profiles.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATHS = {"ieee": HERE / "_build" / "liboracle.so", "fma": HERE / "_build" / "liboracle_fma.so"}

_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)


class OrcA2D(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int), ("nx", ctypes.c_int), ("B", ctypes.c_int), ("M", ctypes.c_int),
        ("nt", ctypes.c_int), ("vp_batched", ctypes.c_int), ("abcn", ctypes.c_int),
        ("pad", ctypes.c_int * 4), ("dt", ctypes.c_float), ("dx", ctypes.c_float),
        ("dz", ctypes.c_float),
        ("az", _F), ("bz", _F), ("dbz", _F), ("ax", _F), ("bx", _F), ("dbx", _F),
        ("lap", _F), ("grad", _F), ("slap", _F), ("slap_c0", ctypes.c_float), ("sgrad", _F),
    ]


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib(variant="ieee"):
    """The oracle library: ``'ieee'`` (the checker) or ``'fma'`` (the same
    source with FMA contraction, for rounding-sensitivity measurements)."""
    if variant not in _libs:
        path = LIB_PATHS[variant]
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        L.orc_a2d_forward.restype = None
        L.orc_a2d_gradient.restype = ctypes.c_int
        L.orc_a2d_strip_cells.restype = ctypes.c_long
        L.orc_a2d_fwd_step.restype = None
        L.orc_a2d_adj_step.restype = None
        L.orc_set_edge_mode.restype = None
        L.orc_set_strip_bf16.restype = None
        L.orc_set_strip_bf16.argtypes = [ctypes.c_int]
        L.orc_bf16_round.restype = ctypes.c_float
        L.orc_bf16_round.argtypes = [ctypes.c_float]
        L.orc_a3d_forward.restype = None
        L.orc_a3d_gradient.restype = ctypes.c_int
        L.orc_e2d_forward.restype = None
        L.orc_e2d_gradient.restype = ctypes.c_int
        _libs[variant] = L
    return _libs[variant]


def _fp(a):
    return a.ctypes.data_as(_F) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class A2DProblem:
    """Everything the C oracle needs for one 2-D acoustic run, on the runtime
    grid.  Keeps the numpy arrays alive for the struct's pointers."""

    def __init__(self, order, shape, dh, dt, nt, abcn=50, max_vel=4500.0, pml_freq=25.0,
                 B=1, variant="ieee"):
        from sweep_amd import pml, scalars
        M = order // 2
        self.M, self.nt, self.B = M, nt, B
        self.abcn, self.dt, self.dh_zyx = abcn, dt, (dh, dh)
        self.variant = variant
        pad = [abcn] * 4
        self.pad = pad
        self.nz = shape[0] + pad[0] + pad[1] + 2 * M
        self.nx = shape[1] + pad[2] + pad[3] + 2 * M
        prof = pml.cpml_profiles_r(pad, [M] * 4, dt, [dh, dh], max_vel, pml_freq,
                                   (self.nz, self.nx))
        self.prof = [_f32(p) for p in prof]
        lap, grad = scalars.laplace_and_gradient_tables(M)
        self.lap, self.grad = _f32(lap), _f32(grad)
        sl = scalars.static_laplace_constants(M)
        sg = scalars.static_gradient_constants(M)
        self.slap = _f32(sl) if sl is not None else _f32(np.zeros(1))
        self.sgrad = _f32(sg) if sg is not None else _f32(np.zeros(1))
        c0 = scalars.static_laplace_c0(M)
        g = OrcA2D()
        g.nz, g.nx, g.B, g.M, g.nt = self.nz, self.nx, B, M, nt
        g.vp_batched, g.abcn = 1, abcn
        for k in range(4):
            g.pad[k] = pad[k]
        g.dt, g.dx, g.dz = dt, dh, dh
        g.az, g.bz, g.dbz, g.ax, g.bx, g.dbx = (_fp(p) for p in self.prof)
        g.lap, g.grad = _fp(self.lap), _fp(self.grad)
        g.slap, g.sgrad = _fp(self.slap), _fp(self.sgrad)
        g.slap_c0 = float(c0) if c0 is not None else 0.0
        self.g = g

    # ---- host-side assembly -------------------------------------------------
    def pad_model(self, vp):
        """(nz_phys, nx_phys) or (B, ...) -> (B, nz, nx) replicate-padded."""
        vp = np.asarray(vp, dtype=np.float32)
        if vp.ndim == 2:
            vp = np.broadcast_to(vp, (self.B,) + vp.shape)
        p = self.pad[0] + self.M, self.pad[1] + self.M, self.pad[2] + self.M, self.pad[3] + self.M
        return _f32(np.pad(vp, ((0, 0), (p[0], p[1]), (p[2], p[3])), mode="edge"))

    def shift(self, loc):
        """Physical (x, z) int coords (B, n, 2) -> runtime int32."""
        loc = np.array(loc, dtype=np.int64).copy()
        loc[..., 0] += self.pad[2] + self.M
        loc[..., 1] += self.pad[0] + self.M
        return np.ascontiguousarray(loc, dtype=np.int32)

    def crop(self, a):
        p = self.pad
        return a[..., p[0] + self.M:self.nz - p[1] - self.M, p[2] + self.M:self.nx - p[3] - self.M]

    # ---- runs ---------------------------------------------------------------
    def forward(self, vp_pad, wav, sloc, rloc):
        B, nt = self.B, self.nt
        nsrc, nrec = sloc.shape[1], rloc.shape[1]
        rec = np.zeros((B, nrec, nt), dtype=np.float32)
        lib(self.variant).orc_a2d_forward(ctypes.byref(self.g), _fp(_f32(vp_pad)), _fp(_f32(wav)), _ip(sloc),
                              nsrc, _ip(rloc), nrec, _fp(rec))
        return rec

    def gradient(self, mode, vp_pad, wav, sloc, rloc, res, illumination=False):
        B, nt = self.B, self.nt
        nsrc, nrec = sloc.shape[1], rloc.shape[1]
        n = self.nz * self.nx
        rec = np.zeros((B, nrec, nt), dtype=np.float32)
        gv = np.zeros((B, self.nz, self.nx), dtype=np.float32)
        gw = np.zeros((B, nsrc, nt), dtype=np.float32)
        S = np.zeros_like(gv) if illumination else None
        R = np.zeros_like(gv) if illumination else None
        rc = lib(self.variant).orc_a2d_gradient(ctypes.byref(self.g), int(mode), _fp(_f32(vp_pad)),
                                    _fp(_f32(wav)), _ip(sloc), nsrc, _ip(rloc), nrec,
                                    _fp(_f32(res)), _fp(rec), _fp(gv), _fp(gw), _fp(S), _fp(R))
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        del n
        out = dict(rec=rec, grad_vp=gv, grad_wavelet=gw)
        if illumination:
            out.update(source_illumination=S, receiver_illumination=R)
        return out


class OrcA3D(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int), ("ny", ctypes.c_int), ("nx", ctypes.c_int), ("B", ctypes.c_int),
        ("M", ctypes.c_int), ("nt", ctypes.c_int), ("vp_batched", ctypes.c_int), ("abcn", ctypes.c_int),
        ("pad", ctypes.c_int * 6), ("dt", ctypes.c_float), ("dx", ctypes.c_float), ("dy", ctypes.c_float),
        ("dz", ctypes.c_float),
        ("az", _F), ("bz", _F), ("dbz", _F), ("ay", _F), ("by", _F), ("dby", _F),
        ("ax", _F), ("bx", _F), ("dbx", _F),
        ("lap", _F), ("grad", _F), ("slap", _F), ("slap_c0", ctypes.c_float), ("sgrad", _F),
    ]


class A3DProblem:
    """The 3-D twin of :class:`A2DProblem`: runtime grid (nz, ny, nx) =
    physical + pad + 2M per axis, coordinates in (x, y, z) order
    (``core/geometry.py:48-56``)."""

    def __init__(self, order, shape, dh, dt, nt, abcn=50, max_vel=4500.0, pml_freq=25.0, B=1,
                 variant="ieee"):
        from sweep_amd import pml, scalars
        M = order // 2
        self.M, self.nt, self.B = M, nt, B
        dh = tuple(dh) if np.ndim(dh) else (dh, dh, dh)
        self.abcn, self.dt, self.dh_zyx = abcn, dt, dh
        self.variant = variant
        pad = [abcn] * 6
        self.pad = pad
        self.nz = shape[0] + pad[0] + pad[1] + 2 * M
        self.ny = shape[1] + pad[2] + pad[3] + 2 * M
        self.nx = shape[2] + pad[4] + pad[5] + 2 * M
        prof = pml.cpml_profiles_r(pad, [M] * 6, dt, list(dh), max_vel, pml_freq, (self.nz, self.ny, self.nx))
        self.prof = [_f32(p) for p in prof]
        lap, grad = scalars.laplace_and_gradient_tables(M)
        self.lap, self.grad = _f32(lap), _f32(grad)
        sl = scalars.static_laplace_constants(M)
        sg = scalars.static_gradient_constants(M)
        self.slap = _f32(sl) if sl is not None else _f32(np.zeros(1))
        self.sgrad = _f32(sg) if sg is not None else _f32(np.zeros(1))
        c0 = scalars.static_laplace_c0(M)
        g = OrcA3D()
        g.nz, g.ny, g.nx, g.B, g.M, g.nt = self.nz, self.ny, self.nx, B, M, nt
        g.vp_batched, g.abcn = 1, abcn
        for k in range(6):
            g.pad[k] = pad[k]
        g.dt, g.dz, g.dy, g.dx = dt, dh[0], dh[1], dh[2]
        (g.az, g.bz, g.dbz, g.ay, g.by, g.dby, g.ax, g.bx, g.dbx) = (_fp(p) for p in self.prof)
        g.lap, g.grad = _fp(self.lap), _fp(self.grad)
        g.slap, g.sgrad = _fp(self.slap), _fp(self.sgrad)
        g.slap_c0 = float(c0) if c0 is not None else 0.0
        self.g = g

    def pad_model(self, vp):
        vp = np.asarray(vp, dtype=np.float32)
        if vp.ndim == 3:
            vp = np.broadcast_to(vp, (self.B,) + vp.shape)
        q = [p + self.M for p in self.pad]
        return _f32(np.pad(vp, ((0, 0), (q[0], q[1]), (q[2], q[3]), (q[4], q[5])), mode="edge"))

    def shift(self, loc):
        """Physical (x, y, z) int coords (B, n, 3) -> runtime int32."""
        loc = np.array(loc, dtype=np.int64).copy()
        loc[..., 0] += self.pad[4] + self.M
        loc[..., 1] += self.pad[2] + self.M
        loc[..., 2] += self.pad[0] + self.M
        return np.ascontiguousarray(loc, dtype=np.int32)

    def crop(self, a):
        p, M = self.pad, self.M
        return a[..., p[0] + M:self.nz - p[1] - M, p[2] + M:self.ny - p[3] - M, p[4] + M:self.nx - p[5] - M]

    def forward(self, vp_pad, wav, sloc, rloc):
        rec = np.zeros((self.B, rloc.shape[1], self.nt), dtype=np.float32)
        lib(self.variant).orc_a3d_forward(ctypes.byref(self.g), _fp(_f32(vp_pad)), _fp(_f32(wav)), _ip(sloc),
                                          sloc.shape[1], _ip(rloc), rloc.shape[1], _fp(rec))
        return rec

    def gradient(self, mode, vp_pad, wav, sloc, rloc, res, illumination=False):
        B, nt = self.B, self.nt
        nsrc, nrec = sloc.shape[1], rloc.shape[1]
        rec = np.zeros((B, nrec, nt), dtype=np.float32)
        gv = np.zeros((B, self.nz, self.ny, self.nx), dtype=np.float32)
        gw = np.zeros((B, nsrc, nt), dtype=np.float32)
        S = np.zeros_like(gv) if illumination else None
        R = np.zeros_like(gv) if illumination else None
        rc = lib(self.variant).orc_a3d_gradient(ctypes.byref(self.g), int(mode), _fp(_f32(vp_pad)), _fp(_f32(wav)),
                                                _ip(sloc), nsrc, _ip(rloc), nrec, _fp(_f32(res)), _fp(rec),
                                                _fp(gv), _fp(gw), _fp(S), _fp(R))
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        out = dict(rec=rec, grad_vp=gv, grad_wavelet=gw)
        if illumination:
            out.update(source_illumination=S, receiver_illumination=R)
        return out


class OrcE2D(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int), ("nx", ctypes.c_int), ("B", ctypes.c_int), ("M", ctypes.c_int),
        ("nt", ctypes.c_int), ("vp_batched", ctypes.c_int), ("pad", ctypes.c_int * 4),
        ("dt", ctypes.c_float), ("dx", ctypes.c_float), ("dz", ctypes.c_float),
        ("az", _F), ("bz", _F), ("azh", _F), ("bzh", _F), ("ax", _F), ("bx", _F), ("axh", _F), ("bxh", _F),
        ("coef", _F), ("nsf", ctypes.c_int), ("nrf", ctypes.c_int), ("src_fields", _I), ("rec_fields", _I),
        ("bs", ctypes.c_int),
    ]


FIELDS = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


class E2DProblem:
    """Everything the elastic oracle needs: runtime grid (nz, nx), cpmls
    profiles [az, bz, azh, bzh, ax, bx, axh, bxh] with fd_pad = M, the
    staggered weights, and the source / receiver field lists."""

    def __init__(self, order, shape, dh, dt, nt, abcn=50, max_vel=4500.0, pml_freq=25.0, B=1,
                 source_fields=("vz",), receiver_fields=("vx", "vz"), variant="ieee", bs=False):
        from sweep_amd import pml, scalars
        M = order // 2
        self.M, self.nt, self.B = M, nt, B
        self.abcn, self.dt, self.dh_zyx = abcn, dt, (dh, dh)
        self.variant = variant
        pad = [abcn] * 4
        self.pad = pad
        self.nz = shape[0] + pad[0] + pad[1] + 2 * M
        self.nx = shape[1] + pad[2] + pad[3] + 2 * M
        prof = pml.cpml_profiles_s(pad, [M] * 4, dt, [dh, dh], max_vel, pml_freq, (self.nz, self.nx))
        self.prof = [_f32(p) for p in prof]
        coef = np.zeros(M + 1, dtype=np.float32)
        coef[:M] = scalars.staggered_grid_coes(M)
        self.grad = coef
        self.sf = np.array([FIELDS[f] for f in source_fields], dtype=np.int32)
        self.rf = np.array([FIELDS[f] for f in receiver_fields], dtype=np.int32)
        g = OrcE2D()
        g.nz, g.nx, g.B, g.M, g.nt, g.vp_batched = self.nz, self.nx, B, M, nt, 1
        for k in range(4):
            g.pad[k] = pad[k]
        g.dt, g.dx, g.dz = dt, dh, dh
        (g.az, g.bz, g.azh, g.bzh, g.ax, g.bx, g.axh, g.bxh) = (_fp(p) for p in self.prof)
        g.coef = _fp(self.grad)
        g.nsf, g.nrf = len(self.sf), len(self.rf)
        g.src_fields, g.rec_fields = _ip(self.sf), _ip(self.rf)
        g.bs = 1 if bs else 0
        self.g = g

    def pad_model(self, m):
        m = np.asarray(m, dtype=np.float32)
        if m.ndim == 2:
            m = np.broadcast_to(m, (self.B,) + m.shape)
        p = [q + self.M for q in self.pad]
        return _f32(np.pad(m, ((0, 0), (p[0], p[1]), (p[2], p[3])), mode="edge"))

    def shift(self, loc):
        loc = np.array(loc, dtype=np.int64).copy()
        loc[..., 0] += self.pad[2] + self.M
        loc[..., 1] += self.pad[0] + self.M
        return np.ascontiguousarray(loc, dtype=np.int32)

    def crop(self, a):
        p, M = self.pad, self.M
        return a[..., p[0] + M:self.nz - p[1] - M, p[2] + M:self.nx - p[3] - M]

    def forward(self, vp, vs, rho, wav, sloc, rloc):
        rec = np.zeros((len(self.rf), self.B, rloc.shape[1], self.nt), dtype=np.float32)
        lib(self.variant).orc_e2d_forward(ctypes.byref(self.g), _fp(_f32(vp)), _fp(_f32(vs)), _fp(_f32(rho)),
                                          _fp(_f32(wav)), _ip(sloc), sloc.shape[1], _ip(rloc), rloc.shape[1],
                                          _fp(rec))
        return rec

    def gradient(self, vp, vs, rho, wav, sloc, rloc, res):
        B, nt = self.B, self.nt
        rec = np.zeros((len(self.rf), B, rloc.shape[1], nt), dtype=np.float32)
        g = [np.zeros((B, self.nz, self.nx), dtype=np.float32) for _ in range(3)]
        rc = lib(self.variant).orc_e2d_gradient(ctypes.byref(self.g), _fp(_f32(vp)), _fp(_f32(vs)),
                                                _fp(_f32(rho)), _fp(_f32(wav)), _ip(sloc), sloc.shape[1],
                                                _ip(rloc), rloc.shape[1], _fp(_f32(res)), _fp(rec), _fp(g[0]),
                                                _fp(g[1]), _fp(g[2]))
        if rc != 0:
            raise MemoryError("oracle allocation failed")
        return dict(rec=rec, grad_vp=g[0], grad_vs=g[1], grad_rho=g[2])


def a2d_from_golden(z, variant="ieee"):
    """Problem + inputs of one golden case (tests/golden/acoustic2d_*.npz)."""
    shape = tuple(int(v) for v in z["shape"])
    src = np.asarray(z["sources"])
    rec = np.asarray(z["receivers"])
    B = src.shape[0]
    P = A2DProblem(int(z["order"]), shape, float(z["dh"]), float(z["dt"]), int(z["nt"]),
                   abcn=int(z["abcn"]), B=B, variant=variant)
    wav = np.broadcast_to(np.asarray(z["wavelet"], dtype=np.float32), (B, 1, int(z["nt"])))
    return P, _f32(wav), P.shift(src[:, None, :]), P.shift(rec)


def num_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
