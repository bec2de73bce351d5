"""``PropTorch``: the propagator API of the reference, backed by libswp_core.

Mirrors ``propagator/torch.py:576-825`` and the compiled backend
``propagator/_c.py`` of the reference for the constant-density acoustic
(2-D/3-D) and isotropic elastic (2-D) equations:

* construction takes ``(equation, shape, dh, dt, device, impl, memory, abcn,
  ...)``; ``impl`` resolves to ``'c'`` on a GPU device and refuses anything
  else (``'eager'`` and the CPU host core are not part of this build);
* ``forward(wavelet, sources, receivers, models=[...])`` accepts the three
  acquisition modes of ``core/validation.py`` and returns the record in the
  canonical ``(B, nt, nrec, nfield)`` layout (``_c.py:111-124``);
* the autograd Function :class:`_Wrapper` owns every buffer (workspace, the
  forward->backward tape, record, gradients), hands them to the core by
  pointer, and refuses a second backward over one graph and
  ``create_graph`` (``_c.py:602-625``); a non-float32 model is a
  ``TypeError`` (``_c.py:384-400``);
* the model is replicate-padded by pad + M (``EdgePadding``) and used shared
  across the shots, or per shot for a ``(B, ...)`` model
  (``_c.py:2110-2155``); the per-shot gradients are summed for a shared model,
  which is what autograd does with the reference's ``repeat``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch
from torch.autograd.function import once_differentiable

from . import _core, geometry, pml, records, scalars
from .equations import Acoustic, Acoustic3D, Elastic, WaveEquation
from .options import BoundarySaving, Ckpt, Full, as_memory_strategy

_CONSTRUCTION_ONLY = ("free_surface", "abcn", "pad", "pml_type", "topography")


def _ptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_float)) if t is not None else None


def _iptr(t):
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctypes.c_int32))


@dataclass
class _Call:
    """Everything one compiled call needs; the tensors stay referenced here
    for the call's lifetime."""
    eq: WaveEquation
    names: tuple
    problem: _core.SwpProblem
    keep: list = field(default_factory=list)
    shared_model: bool = True
    nrec: int = 0
    nfield: int = 1
    device: torch.device = None
    memory: int = 0
    requires_backward: bool = False
    arena: object = None
    illumination: bool = False


class _Arena:
    """Device buffers of one call shape kept from call to call (workspace,
    tape, raw record, and staging copies of the inputs and outputs), so the
    core sees the same addresses every time and replays the HIP graph it
    captured of the call's time loop (SWP_FLAG_GRAPH).  A forward that keeps
    a tape for its backward holds the arena until that backward ran or its
    autograd graph was dropped."""

    def __init__(self, dev):
        self.dev = dev
        self.bufs = {}
        self.busy = False

    def get(self, name, shape, dtype=torch.float32):
        t = self.bufs.get(name)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(tuple(shape), dtype=dtype, device=self.dev)
            self.bufs[name] = t
        return t

    def get_bytes(self, name, n):
        """A uint8 buffer of at least n bytes (grown, never shrunk)."""
        t = self.bufs.get(name)
        if t is None or t.numel() < n:
            self.bufs.pop(name, None)
            t = torch.empty(max(int(n), 1), dtype=torch.uint8, device=self.dev)
            self.bufs[name] = t
        return t

    def stage(self, name, src):
        t = self.get(name, src.shape, src.dtype)
        t.copy_(src)
        return t

    def release(self):
        self.busy = False


def _release_arena(arena):
    arena.release()


class _Wrapper(torch.autograd.Function):

    @staticmethod
    def forward(ctx, call: _Call, wavelet, *models):
        P = call.problem
        dev = call.device
        A = call.arena
        P.memory = call.memory if call.requires_backward else _core.MEM_NONE
        B, nt = P.B, P.nt
        rshape = (call.nfield, B, call.nrec, nt) if call.nfield > 1 else (B, call.nrec, nt)
        if A is not None:
            # the arena's fixed buffers: stage the inputs, replay the graph
            P.flags |= _core.FLAG_GRAPH
            P.source = _ptr(A.stage("wavelet", wavelet))
            for k, m in enumerate(models):
                P.models[k] = _ptr(A.stage(f"model{k}", m))
            P.sources_loc = _iptr(A.stage("sources_loc", call.keep[0]))
            P.receivers_loc = _iptr(A.stage("receivers_loc", call.keep[1]))
            ws, tp = _core.query(call.names[0], P)
            workspace = A.get_bytes("workspace", ws)
            tape = A.get_bytes("tape", tp) if P.memory != _core.MEM_NONE else None
            record = A.get("record", rshape)
            record.zero_()
            if tape is not None:
                A.busy = True
        else:
            P.source = _ptr(wavelet)
            for k, m in enumerate(models):
                P.models[k] = _ptr(m)
            ws, tp = _core.query(call.names[0], P)
            workspace = torch.empty(ws, dtype=torch.uint8, device=dev)
            tape = torch.empty(tp, dtype=torch.uint8, device=dev) if P.memory != _core.MEM_NONE else None
            record = torch.zeros(rshape, dtype=torch.float32, device=dev)
        P.workspace, P.workspace_bytes = workspace.data_ptr(), workspace.numel()
        P.tape, P.tape_bytes = (tape.data_ptr(), tape.numel()) if tape is not None else (None, 0)
        io = _core.SwpForwardIO(record=_ptr(record))
        call.eq._C()[0](P, io, dev)
        ctx.call = call
        ctx.tape = tape
        ctx.used = False
        if A is not None and tape is not None:
            # the tape's arena is free again once this graph is gone
            import weakref
            weakref.finalize(ctx, _release_arena, A)
        ctx.save_for_backward(wavelet, *models)
        return records.to_canonical(record)

    @staticmethod
    @once_differentiable
    def backward(ctx, grad):
        if ctx.used:
            raise RuntimeError("impl='c' backward was already run over this graph; a second backward "
                               "(retain_graph=True) is refused: the tape was consumed.")
        ctx.used = True
        call = ctx.call
        saved = ctx.saved_tensors
        wavelet, models = saved[0], saved[1:]
        P = call.problem
        dev = call.device
        if grad is None:
            return (None, None) + (None,) * len(models)
        adj = records.to_raw(grad.to(torch.float32), 4 if call.nfield > 1 else 3)
        entry = {_core.MEM_FULL: 1, _core.MEM_BOUNDARY: 2, _core.MEM_CKPT: 3, _core.MEM_BINOMIAL: 4}[P.memory]
        name = call.names[entry]
        if call.illumination:
            P.flags |= _core.FLAG_ILLUMINATION   # before the query: the workspace depends on it
        A = call.arena
        ws, _ = _core.query(name, P)
        grid = (P.B,) + ((P.nz, P.ny, P.nx) if P.ndim == 3 else (P.nz, P.nx))
        want_w = ctx.needs_input_grad[1]
        want_m = [ctx.needs_input_grad[2 + k] for k in range(len(models))]

        def out_buf(name, shape):
            if A is None:
                return torch.zeros(shape, dtype=torch.float32, device=dev)
            t = A.get(name, shape)
            t.zero_()
            return t

        workspace = A.get_bytes("workspace", ws) if A is not None else torch.empty(ws, dtype=torch.uint8, device=dev)
        P.workspace, P.workspace_bytes = workspace.data_ptr(), workspace.numel()
        g_w = out_buf("grad_wavelet", wavelet.shape) if want_w else None
        g_m = [out_buf(f"grad_model{k}", grid) if want_m[k] else None for k in range(len(models))]
        if A is not None:
            adj = A.stage("adjoint_source", adj)
        io = _core.SwpBackwardIO()
        io.adjoint_source = _ptr(adj)
        io.grad_source = _ptr(g_w)
        for k, g in enumerate(g_m):
            io.grad_models[k] = _ptr(g)
        S = R = None
        if call.illumination:
            S = out_buf("illum_source", grid)
            R = out_buf("illum_receiver", grid)
            io.illum_source, io.illum_receiver = _ptr(S), _ptr(R)
        P.tape, P.tape_bytes = ctx.tape.data_ptr(), ctx.tape.numel()
        call.eq._C()[entry](P, io, dev)
        ctx.tape = None
        if A is not None:
            # the caller gets tensors of its own; the arena's are reused
            g_w = g_w.clone() if g_w is not None else None
            g_m = [g.clone() if g is not None else None for g in g_m]
            S, R = (S.clone(), R.clone()) if S is not None else (None, None)
            A.release()
        if call.illumination:
            call.illum_out = (S, R)
        out_m = []
        for g in g_m:
            if g is None:
                out_m.append(None)
            elif call.shared_model:
                out_m.append(g.sum(0))
            else:
                out_m.append(g)
        return (None, g_w) + tuple(out_m)


class PropTorch(torch.nn.Module):
    """Finite-difference propagator on one MI355X (``impl='c'``)."""

    def __init__(self, equation, shape, dh=10.0, dt=0.002, device=None, dev=None, impl=None,
                 memory=None, abcn=50, free_surface=False, nt=-1, B=1, source_type=None,
                 receiver_type=None, compute_illumination=False, use_ckpt=None, topography=None,
                 backend=None, cuda_options=None, eager_options=None, graphs=True, **kwargs):
        super().__init__()
        if not isinstance(equation, WaveEquation):
            raise TypeError("equation must be an Acoustic / Acoustic3D / Elastic instance")
        device = device if device is not None else dev
        device = torch.device(device if device is not None else
                              (equation.device if equation.device is not None else "cuda"))
        if impl in (None, "auto", "c"):
            if device.type != "cuda":
                raise RuntimeError(
                    "impl='c' runs on the GPU only in this build (the CPU host core and the eager "
                    "PyTorch path are not part of it); pass device='cuda'.")
            impl = "c"
        elif impl == "eager":
            raise NotImplementedError("impl='eager' is not part of this build; use impl='c'.")
        else:
            raise ValueError(f"impl must be 'c', 'auto' or None, got {impl!r}")
        if free_surface or topography is not None:
            raise NotImplementedError("free surface / topography are not part of this build (CPML on every face).")
        if eager_options is not None:
            raise ValueError("eager_options applies to impl='eager', which this build does not have")
        if use_ckpt is True and memory is None:
            memory = Ckpt()
        if memory is not None:
            memory = as_memory_strategy(memory)
        else:
            memory = BoundarySaving()   # the compiled path's default (propagator/torch.py:594-599)
        if isinstance(memory, BoundarySaving):
            memory.check_supported()
            if isinstance(equation, Elastic) and memory.storage_dtype != "fp32":
                raise NotImplementedError("Elastic boundary saving keeps float32 strips in this build "
                                          "(BoundarySaving(storage_dtype='fp32')).")
        if isinstance(memory, Ckpt) and memory.storage != "gpu":
            raise NotImplementedError("Ckpt(storage='cpu') is not part of this build.")
        self.__dict__["_built"] = False
        self.equation = equation
        self.impl = impl
        self.dev = device
        self.memory = memory
        self.ndim = equation.ndim
        if len(shape) != self.ndim:
            raise ValueError(f"shape {tuple(shape)} does not match the equation's {self.ndim} dims")
        self._shape_phys = tuple(int(s) for s in shape)
        if isinstance(abcn, (list, tuple)):
            pad = tuple(int(p) for p in abcn)
            if len(pad) != 2 * self.ndim:
                raise ValueError(f"abcn per edge needs {2 * self.ndim} values")
        else:
            pad = (int(abcn),) * (2 * self.ndim)
        self.abcn = max(pad)
        self.pad = pad
        self.pml_pad = pad
        dh = (float(dh),) * self.ndim if np.isscalar(dh) else tuple(float(v) for v in dh)
        self._grid_spacing = dh
        self._dt = float(dt)
        self.nt = nt
        self.B = B
        self.source_type = list(source_type) if source_type else list(equation.default_source_fields)
        self.receiver_type = list(receiver_type) if receiver_type else list(equation.default_receiver_fields)
        self.compute_illumination = bool(compute_illumination)
        self.source_illumination = None
        self.receiver_illumination = None
        self.pml_type = equation.default_pml_type
        M = equation.M
        self.shape = tuple(self._shape_phys[a] + pad[2 * a] + pad[2 * a + 1] for a in range(self.ndim))
        self.shape_cuda = tuple(s + 2 * M for s in self.shape)
        self._abc_cache = {}
        # graphs=True: calls run on per-shape arenas of fixed buffers and the
        # core replays each call's captured time-loop graph (an extension:
        # the reference launches step by step); False allocates per call
        self.graphs = bool(graphs)
        self._arenas = {}
        self.funcs = equation._C()
        self.__dict__["_built"] = True

    def __setattr__(self, name, value):
        if self.__dict__.get("_built") and name in _CONSTRUCTION_ONLY:
            raise AttributeError(f"{name!r} is fixed at construction; build a new PropTorch instead")
        super().__setattr__(name, value)

    @property
    def dt(self):
        return self._dt

    @property
    def dh(self):
        return self._grid_spacing

    # ---- geometry ---------------------------------------------------------
    def _profiles(self, max_vel, pml_freq):
        key = (max_vel, pml_freq)
        if key in self._abc_cache:
            return self._abc_cache[key]
        M = self.equation.M
        fd_pad = [M, M] * self.ndim
        # order in shape axes (z[, y], x); pml pad per face in that order
        if self.pml_type == "cpmlr":
            prof = pml.cpml_profiles_r(list(self.pml_pad), fd_pad, self._dt, list(self._grid_spacing),
                                       max_vel, pml_freq, self.shape_cuda)
        else:
            prof = pml.cpml_profiles_s(list(self.pml_pad), fd_pad, self._dt, list(self._grid_spacing),
                                       max_vel, pml_freq, self.shape_cuda)
        tens = [torch.from_numpy(np.ascontiguousarray(p, dtype=np.float32)).to(self.dev) for p in prof]
        lap, grad = scalars.laplace_and_gradient_tables(M)
        if isinstance(self.equation, Elastic):
            grad = np.zeros(M + 1, dtype=np.float32)
            grad[:M] = scalars.staggered_grid_coes(M)
        self._abc_cache[key] = (prof, tens, np.ascontiguousarray(lap), np.ascontiguousarray(grad))
        return self._abc_cache[key]

    def _runtime_coords(self, loc):
        off = geometry.runtime_coord_offset(self.pad, self.equation.M, self.ndim)
        if isinstance(loc, torch.Tensor):   # coordinates may live on the GPU
            loc = loc.detach().cpu().numpy()
        loc = np.array(loc, dtype=np.int64, copy=True)
        for a in range(self.ndim):
            loc[..., a] += off[a]
        return loc

    _FIELD_NAMES = {"vx": 0, "velocity_x": 0, "vz": 1, "velocity_z": 1, "sxx": 2, "stress_xx": 2, "szz": 3,
                    "stress_zz": 3, "sxz": 4, "stress_xz": 4, "shear_xz": 4}

    def _field_indices(self):
        """Source / receiver wavefield indices (elastic: vx, vz, sxx, szz, sxz;
        acoustic: the one pressure field)."""
        if not isinstance(self.equation, Elastic):
            return [0], [0]
        out = []
        for kind, names in (("source", self.source_type), ("receiver", self.receiver_type)):
            idx = []
            for name in names:
                if name not in self._FIELD_NAMES:
                    raise ValueError(f"unknown elastic {kind} field {name!r}; expected one of vx, vz, sxx, szz, sxz")
                i = self._FIELD_NAMES[name]
                if kind == "receiver" and i == 4:
                    raise ValueError("sxz does not support receivers (equations/elastic.py FIELD_SPECS)")
                idx.append(i)
            if not 1 <= len(idx) <= 5:
                raise ValueError(f"need 1..5 {kind} fields, got {len(idx)}")
            out.append(idx)
        return out[0], out[1]

    def crop(self, data):
        """Runtime/padded-grid array -> physical box."""
        return geometry.crop_physical(data, self.pad, self.equation.M, self.ndim)

    # ---- call -------------------------------------------------------------
    def forward(self, wavelet, sources, receivers, models=None, adj=False, return_wavefield=False,
                **kwargs):
        if adj or return_wavefield:
            raise NotImplementedError("adj / return_wavefield are not part of this build")
        unknown = set(kwargs) - {"max_vel", "pml_freq", "fd_pad", "shape"}
        if unknown:
            raise TypeError(f"forward() got unexpected keyword arguments: {sorted(unknown)}; "
                            "recognised extras are ['fd_pad', 'max_vel', 'pml_freq', 'shape']")
        eq = self.equation
        ndim = self.ndim
        M = eq.M
        mode, B, nsrc, nrec, _ = geometry.normalize_io(wavelet, sources, receivers, ndim)
        dev = self.dev
        models = list(models if models is not None else [])
        if len(models) != len(eq.model_names):
            raise ValueError(f"{type(eq).__name__} takes models {list(eq.model_names)}, got {len(models)}")
        for name, m in zip(eq.model_names, models):
            if not isinstance(m, torch.Tensor) or m.dtype != torch.float32:
                raise TypeError(f"impl='c' computes in float32: model {name!r} must be a float32 "
                                f"torch.Tensor, got {getattr(m, 'dtype', type(m))}")
        # wavelet -> (B, nsrc, nt) float32, autograd-tracked back to the caller's shape
        w = wavelet if isinstance(wavelet, torch.Tensor) else torch.from_numpy(np.asarray(wavelet))
        w = w.to(dev, dtype=torch.float32)
        nt = int(w.shape[-1])
        if mode == "A1":
            w = w[None, None, :].expand(B, 1, nt)
        elif mode == "A2":
            w = w[:, None, :]
        else:
            w = w[None, None, :].expand(1, nsrc, nt) if w.ndim == 1 else w[None, :, :]
        w = w.contiguous()
        src = self._runtime_coords(sources)
        if mode in ("A1", "A2"):
            src = src[:, None, :]
        rec = self._runtime_coords(receivers)
        src_t = torch.from_numpy(np.ascontiguousarray(src, dtype=np.int32)).to(dev)
        rec_t = torch.from_numpy(np.ascontiguousarray(rec, dtype=np.int32)).to(dev)
        prof, prof_t, lap, grad = self._profiles(kwargs.get("max_vel", 4500.0), kwargs.get("pml_freq", 25.0))
        torch_pad = geometry.torch_pad_order(self.pad, ndim, M)
        padded = []
        shared = None
        for m in models:
            m = m.to(dev)
            if m.ndim == ndim:
                s = True
            elif m.ndim == ndim + 1:
                if m.shape[0] != B:
                    raise ValueError(f"Per-shot model batch ({int(m.shape[0])}) must equal the shot batch "
                                     f"size ({B}); provide exactly one model per shot.")
                s = False
            else:
                raise ValueError(f"Model tensor has {m.ndim} dims; expected {ndim} or {ndim + 1}.")
            if shared is None:
                shared = s
            elif shared != s:
                raise ValueError("models must be all shared or all per-shot")
            padded.append(geometry.EdgePadding.apply(m, torch_pad).contiguous())
        grid = self.shape_cuda
        for p in padded:
            if tuple(p.shape[-ndim:]) != grid:
                raise ValueError(f"model shape {tuple(p.shape)} does not pad to the runtime grid {grid}")

        P = _core.SwpProblem()
        P.ndim = ndim
        if ndim == 2:
            P.nz, P.ny, P.nx = grid[0], 1, grid[1]
        else:
            P.nz, P.ny, P.nx = grid
        P.B, P.M, P.nt, P.abcn = B, M, nt, self.abcn
        pads6 = [0] * 6
        if ndim == 2:
            pads6[0], pads6[1], pads6[4], pads6[5] = self.pad
        else:
            pads6[:] = list(self.pad)
        for k in range(6):
            P.pad[k] = pads6[k]
        P.dt = self._dt
        hs = self._grid_spacing
        P.dz, P.dy, P.dx = (hs[0], 0.0, hs[1]) if ndim == 2 else hs
        P.lap_coes = lap.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        P.grad_coes = grad.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        for k, t in enumerate(prof_t):
            P.pml[k] = _ptr(t)
        P.model_batched = 0 if shared else 1
        P.nsrc, P.nrec = nsrc, nrec
        sf, rf = self._field_indices()
        P.nsrc_fields, P.nrec_fields = len(sf), len(rf)
        for k, v in enumerate(sf):
            P.src_fields[k] = v
        for k, v in enumerate(rf):
            P.rec_fields[k] = v
        P.sources_loc = _iptr(src_t)
        P.receivers_loc = _iptr(rec_t)
        mem = self.memory
        if isinstance(mem, Full):
            P.memory = _core.MEM_FULL
        elif isinstance(mem, BoundarySaving):
            P.memory = _core.MEM_BOUNDARY
            if mem.storage_dtype == "bf16":
                P.flags |= _core.FLAG_BF16_STRIPS
        elif mem.mode == "recursive":
            # a fixed budget of `count` snapshot slots, reversed on a binomial
            # schedule (the reference bisects recursively, eq_driver.cuh:1055-1259)
            P.memory = _core.MEM_BINOMIAL
            P.ckpt_interval = max(1, min(int(mem.count), max(nt, 1)))
        else:
            P.memory = _core.MEM_CKPT
            P.ckpt_interval = mem.interval(nt)
        requires = torch.is_grad_enabled() and (w.requires_grad or any(p.requires_grad for p in padded))
        if requires and w.requires_grad and isinstance(eq, Elastic):
            # the compiled elastic backward returns the model gradients only (_c.py:2368-2374)
            raise NotImplementedError(
                f"impl='c' {type(eq).__name__} does not return the wavelet gradient (its compiled backward "
                "returns the model gradients only). Detach the wavelet.")
        call = _Call(eq=eq, names=tuple(f"{eq.C_NAME}_{op}" for op in
                                        ("forward", "backward", "backward_bs", "backward_ckpt",
                                         "backward_recursive_ckpt")),
                     problem=P, keep=[src_t, rec_t, prof_t, lap, grad], shared_model=bool(shared),
                     nrec=nrec, nfield=len(self.receiver_type) if isinstance(eq, Elastic) else 1,
                     device=dev, memory=int(P.memory), requires_backward=bool(requires),
                     illumination=self.compute_illumination)
        if self.graphs:
            call.arena = self._arena_for((B, nsrc, nrec, nt, bool(shared), int(P.memory), int(P.flags),
                                          int(P.ckpt_interval), call.requires_backward, len(models)))
        self._last_call = call
        out = _Wrapper.apply(call, w, *padded)
        return out

    _ARENAS_PER_SHAPE = 2

    def _arena_for(self, key):
        """A free arena for calls of this shape, or None (per-call buffers)
        when every arena of the shape holds a tape that is still waiting
        for its backward."""
        pool = self._arenas.setdefault(key, [])
        for a in pool:
            if not a.busy:
                return a
        if len(pool) < self._ARENAS_PER_SHAPE:
            pool.append(_Arena(self.dev))
            return pool[-1]
        return None

    def collect_illumination(self):
        """Source (sum u_tt^2) and receiver (sum lambda^2) illumination of the
        last backward, summed over shots and cropped to the physical box."""
        call = getattr(self, "_last_call", None)
        if call is None or getattr(call, "illum_out", None) is None:
            return None, None
        S, R = call.illum_out
        self.source_illumination = self.crop(S.sum(0))
        self.receiver_illumination = self.crop(R.sum(0))
        return self.source_illumination, self.receiver_illumination
