"""Equation classes with the reference's plugin slot.

Each class declares ``C_NAME``; :meth:`WaveEquation._C` resolves the five
compiled callables ``<C_NAME>_{forward, backward, backward_bs, backward_ckpt,
backward_recursive_ckpt}`` of ``libswp_core`` exactly as the reference's
``equations/base.py:295-324`` resolves ``sweep._C.<C_NAME>_*``.  Each callable
is ``fn(problem, io, device) -> None``.
"""
from __future__ import annotations

from . import _core

_OPS = ("forward", "backward", "backward_bs", "backward_ckpt", "backward_recursive_ckpt")


class _Entry:
    def __init__(self, name):
        self.name = name

    def __call__(self, problem, io, device=None):
        _core.call(self.name, problem, io, device)

    def __repr__(self):
        return f"<libswp_core entry {self.name}>"


class WaveEquation:
    C_NAME: str | None = None
    NDIM = 2
    MODEL_NAMES: tuple = ()
    default_pml_type = "cpmlr"
    default_source_fields = ("h1",)
    default_receiver_fields = ("h1",)

    def __init__(self, spatial_order=4, device="cpu", backend="torch", dim=None):
        if spatial_order < 2 or spatial_order % 2:
            raise ValueError(f"spatial_order must be an even integer >= 2, got {spatial_order}")
        if backend != "torch":
            raise NotImplementedError("only backend='torch' is built (the JAX backend is out of scope)")
        self.so = int(spatial_order)
        self.device = device
        self.backend = backend
        self.ndim = self.NDIM if dim is None else int(dim)

    @property
    def M(self):
        return self.so // 2

    @property
    def model_names(self):
        return self.MODEL_NAMES

    @classmethod
    def _C(cls):
        """The compiled 5-tuple for this equation (forward, backward,
        backward_bs, backward_ckpt, backward_recursive_ckpt)."""
        if cls.C_NAME is None:
            raise NotImplementedError(f"{cls.__name__} has no compiled core")
        return tuple(_Entry(f"{cls.C_NAME}_{op}") for op in _OPS)


class Acoustic(WaveEquation):
    """Second-order 2-D constant-density acoustic equation with cpmlr CPML
    (reference: ``equations/acoustic.py:47-207``)."""
    C_NAME = "acoustic2d"
    NDIM = 2
    MODEL_NAMES = ("vp",)


class Acoustic3D(WaveEquation):
    """Second-order 3-D constant-density acoustic equation with cpmlr CPML
    (reference: ``equations/acoustic3d.py:49-185``)."""
    C_NAME = "acoustic3d"
    NDIM = 3
    MODEL_NAMES = ("vp",)

    def __init__(self, spatial_order=4, device="cpu", backend="torch", dim=3):
        super().__init__(spatial_order, device, backend, dim)


class Elastic(WaveEquation):
    """Isotropic 2-D velocity-stress elastic equation on a staggered grid with
    cpmls CPML (reference: ``equations/elastic.py:19-310``)."""
    C_NAME = "elastic2d"
    NDIM = 2
    MODEL_NAMES = ("vp", "vs", "rho")
    default_pml_type = "cpmls"
    default_source_fields = ("vz",)
    default_receiver_fields = ("vx", "vz")
