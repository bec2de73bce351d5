"""Host-side logic of the propagator (CPU): IO modes, coordinate maps,
padding, options, refusals.  Mirrors the reference's unit tests of
core/validation.py (test_source_io_contract.py) and core/geometry.py."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from sweep_amd import geometry
from sweep_amd.equations import Acoustic, Acoustic3D, Elastic
from sweep_amd.options import BoundarySaving, Ckpt, Full, as_memory_strategy


def test_io_modes():
    assert geometry.normalize_io(np.zeros(10), np.zeros((3, 2)), np.zeros((3, 5, 2)), 2)[:3] == ("A1", 3, 1)
    assert geometry.normalize_io(np.zeros((3, 10)), np.zeros((3, 2)), np.zeros((3, 5, 2)), 2)[:3] == ("A2", 3, 1)
    assert geometry.normalize_io(np.zeros(10), np.zeros((1, 4, 2)), np.zeros((1, 5, 2)), 2)[:3] == ("B", 1, 4)
    assert geometry.normalize_io(np.zeros((4, 10)), np.zeros((1, 4, 3)), np.zeros((1, 5, 3)), 3)[:3] == ("B", 1, 4)


@pytest.mark.parametrize("w,s,r", [
    ((10,), (3, 2), (5, 2)),          # receivers not 3-D
    ((10,), (3, 2), (2, 5, 2)),       # batch mismatch
    ((2, 10), (3, 2), (3, 5, 2)),     # per-shot wavelet count
    ((10,), (2, 4, 2), (2, 5, 2)),    # encoding needs batch 1
    ((3, 10), (1, 4, 2), (1, 5, 2)),  # encoded wavelet count
    ((10,), (3,), (3, 5, 2)),         # sources rank
])
def test_io_refusals(w, s, r):
    with pytest.raises(ValueError):
        geometry.normalize_io(np.zeros(w), np.zeros(s), np.zeros(r), 2)


def test_runtime_coordinate_map():
    # (x, z) order: x moves by its own low pad + M, z by its own
    pad = (7, 9, 11, 13)   # z_lo, z_hi, x_lo, x_hi
    assert geometry.runtime_coord_offset(pad, 4, 2) == (11 + 4, 7 + 4)
    pad3 = (1, 2, 3, 4, 5, 6)   # z, y, x
    assert geometry.runtime_coord_offset(pad3, 2, 3) == (5 + 2, 3 + 2, 1 + 2)
    assert geometry.runtime_shape((96, 128), (50,) * 4, 2) == (200, 232)
    assert geometry.runtime_shape((384, 1152), (50,) * 4, 4) == (492, 1260)
    assert geometry.runtime_shape((256, 512, 512), (50,) * 6, 4) == (364, 620, 620)


def test_edge_padding_replicates_and_crops():
    x = torch.arange(12, dtype=torch.float32).reshape(3, 4).requires_grad_()
    pad = geometry.torch_pad_order((1, 2, 3, 1), 2, 0)
    y = geometry.EdgePadding.apply(x, pad)
    assert y.shape == (3 + 1 + 2, 4 + 3 + 1)
    assert torch.equal(y[1:4, 3:7], x.detach())
    assert torch.equal(y[0, 3:7], x.detach()[0]) and torch.equal(y[:, 0], y[:, 3])
    g = torch.ones_like(y)
    (y * g).sum().backward()
    assert torch.equal(x.grad, torch.ones_like(x))   # cropped, not folded into the edge


def test_crop_physical():
    a = np.arange(20 * 30).reshape(20, 30)
    c = geometry.crop_physical(a, (2, 3, 4, 5), 1, 2)
    assert c.shape == (20 - 2 - 3 - 2, 30 - 4 - 5 - 2)
    assert c[0, 0] == a[3, 5]


def test_ricker_closed_form():
    t = np.linspace(-0.2, 0.2, 101)
    r = geometry.ricker(t, 10.0)
    assert r.dtype == np.float32
    assert r[50] == 1.0
    a = (np.pi * 10 * t) ** 2
    assert np.allclose(r, (1 - 2 * a) * np.exp(-a), atol=1e-7)


def test_options_validation():
    assert isinstance(as_memory_strategy({"kind": "full"}), Full)
    assert isinstance(as_memory_strategy({"strategy": "boundary", "boundary": {"storage": "gpu"}}), BoundarySaving)
    with pytest.raises(ValueError):
        Ckpt(mode="chunk", chunks=0)
    with pytest.raises(ValueError):
        Ckpt(mode="recursive", count=0)
    with pytest.raises(ValueError):
        BoundarySaving(storage="tape")
    with pytest.raises(NotImplementedError):
        BoundarySaving(storage="cpu").check_supported()
    assert Ckpt(chunks=7).interval(400) == 58
    assert Ckpt(mode="recursive", count=5).interval(100) == 20


def test_equation_plugin_slot():
    for cls, name in ((Acoustic, "acoustic2d"), (Acoustic3D, "acoustic3d"), (Elastic, "elastic2d")):
        funcs = cls._C()
        assert [f.name for f in funcs] == [f"{name}_{op}" for op in
                                           ("forward", "backward", "backward_bs", "backward_ckpt",
                                            "backward_recursive_ckpt")]
    with pytest.raises(ValueError):
        Acoustic(spatial_order=5)
    assert Acoustic(spatial_order=8).M == 4 and Acoustic3D().ndim == 3


def test_propagator_refuses_cpu_and_eager():
    from sweep_amd.propagator import PropTorch
    with pytest.raises(RuntimeError, match="GPU only"):
        PropTorch(Acoustic(device="cpu"), shape=(20, 20), device="cpu")
    with pytest.raises(NotImplementedError):
        PropTorch(Acoustic(device="cpu"), shape=(20, 20), device="cpu", impl="eager")
