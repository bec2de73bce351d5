"""The C oracle against the reference's compiled acoustic and elastic paths.

Fixtures ``golden/refcore{2,3}d_*.npz`` and ``golden/refcore_e2d_*.npz`` hold
the reference CPU host core's full-storage outputs (written by
``golden/gen_refcore.py``); where
``oracle/_ref`` is built the same comparison also runs live on fresh inputs.

Bars (relative L2, full storage): 3-D 1e-8 -- records and model gradients
come out bit-exact or within a few ulps, the oracle and the reference core
evaluating the same float operations in the same order, both without FMA
contraction; 2-D 1e-6 for records and 2e-6 for gradients (order 2 is
bit-exact; the reference 2-D core sums some Laplacian terms in another
order).  The oracle's boundary-saving mode is held to 2e-5 against the
reference's full mode.

Wavelet gradient, 3-D: the reference core accumulates it for shot 0 only.
``accumulate_source_grad_3d`` (csrc/cuda/equations/acoustic3d/kernels.cu:256-257)
reads the shot from blockIdx.y and the source from blockIdx.x, but its launch
(``fdtd::Geom::make``, csrc/cuda/launch/config.h:65-69) puts the shot on
grid.x, so with B > 1 every shot b > 0 keeps a zero wavelet gradient.  The
oracle computes every shot; shot 0 is compared here, and the rest is pinned
by the adjoint identity <F w, r> = <w, F^T r>.
"""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_l2

oracle = pytest.importorskip("oracle.oracle")
from oracle import refcore  # noqa: E402

CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore[23]d_o*.npz")))
BS2D_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore2d_bs_*.npz")))
BS3D_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore3d_bs_*.npz")))


def problem(z):
    ndim = len(z["shape"])
    cls = oracle.A2DProblem if ndim == 2 else oracle.A3DProblem
    B = int(z["B"])
    P = cls(int(z["order"]), tuple(int(v) for v in z["shape"]), float(z["dh"]), float(z["dt"]), int(z["nt"]),
            abcn=int(z["abcn"]), B=B)
    nsrc = z["sources"].shape[1]
    wav = np.broadcast_to(np.asarray(z["wavelet"], np.float32), (B, nsrc, int(z["nt"]))).copy()
    return P, P.pad_model(z["vp"]), wav, P.shift(z["sources"]), P.shift(z["receivers"])


def test_fixtures_present():
    assert any(c.startswith("refcore2d") for c in CASES) and any(c.startswith("refcore3d") for c in CASES)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", [0, 1], ids=["full", "bs"])
def test_oracle_matches_reference_core(case, mode):
    z = np.load(GOLDEN / f"{case}.npz")
    P, vp, wav, s, r = problem(z)
    o = P.gradient(mode, vp, wav, s, r, z["residual"])
    three = case.startswith("refcore3d")
    # full mode: the same arithmetic; boundary saving adds the reconstruction's
    # own rounding (it is compared against the reference's full mode)
    bar_rec = 1e-8 if three else 1e-6
    bar_grad = (1e-8 if three else 2e-6) if mode == 0 else 2e-5
    assert rel_l2(o["rec"], z["rec"]) <= bar_rec
    assert rel_l2(o["grad_vp"], z["grad_vp"]) <= bar_grad
    if three:
        assert rel_l2(o["grad_wavelet"][0], z["grad_wavelet"][0]) <= bar_grad
        assert np.all(z["grad_wavelet"][1:] == 0)     # the reference's shot-0-only accumulation
        assert np.abs(o["grad_wavelet"][1:]).max() > 0
    else:
        assert rel_l2(o["grad_wavelet"], z["grad_wavelet"]) <= bar_grad


@pytest.mark.parametrize("case", BS2D_CASES)
def test_oracle_boundary_saving_matches_reference_core(case):
    """The oracle's boundary-saving backward against the reference's own
    (backward_bs_core through oracle/ref_harness.cpp): both rebuild the
    forward wavefield from the saved strips; measured <= 5.6e-6."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, vp, wav, s, r = problem(z)
    o = P.gradient(1, vp, wav, s, r, z["residual"])
    assert rel_l2(o["rec"], z["rec"]) <= 1e-6
    assert rel_l2(o["grad_vp"], z["grad_vp"]) <= 1e-5
    assert rel_l2(o["grad_wavelet"], z["grad_wavelet"]) <= 1e-5


@pytest.mark.parametrize("case", BS3D_CASES)
def test_oracle3d_boundary_saving_matches_reference_core(case):
    """3-D boundary saving against the reference's backward_bs_core: the
    same arithmetic in the same order (bit-exact at orders 2 and 4, 1e-7 at
    order 8); bar 1e-6."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, vp, wav, s, r = problem(z)
    o = P.gradient(1, vp, wav, s, r, z["residual"])
    assert rel_l2(o["rec"], z["rec"]) <= 1e-6
    assert rel_l2(o["grad_vp"], z["grad_vp"]) <= 1e-6
    assert rel_l2(o["grad_wavelet"][0], z["grad_wavelet"][0]) <= 1e-6


ILLUM_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore[23]d_illum_*.npz")))


@pytest.mark.parametrize("case", ILLUM_CASES)
def test_oracle_illumination_matches_reference_core(case):
    """Source (sum u_tt^2) and receiver (sum lambda^2) illumination against
    the reference's compute_illumination, full storage and boundary saving;
    measured <= 1.6e-8."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, vp, wav, s, r = problem(z)
    o = P.gradient(1 if "_bs_" in case else 0, vp, wav, s, r, z["residual"], illumination=True)
    assert len(ILLUM_CASES) >= 4
    assert rel_l2(o["source_illumination"], z["illum_s"]) <= 1e-6
    assert rel_l2(o["receiver_illumination"], z["illum_r"]) <= 1e-6


@pytest.mark.parametrize("case", [c for c in CASES if c.startswith("refcore3d")])
def test_oracle3d_wavelet_adjoint_identity(case):
    """grad_wavelet = F^T r for the linear map w -> record: <F w, r> = <w, F^T r>."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, vp, wav, s, r = problem(z)
    res = z["residual"]
    o = P.gradient(0, vp, wav, s, r, res)
    lhs = np.sum(o["rec"].astype(np.float64) * res, axis=(1, 2))
    rhs = np.sum(wav.astype(np.float64) * o["grad_wavelet"], axis=(1, 2))
    np.testing.assert_allclose(rhs, lhs, rtol=2e-5)


@pytest.mark.skipif(not refcore.available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("ndim,order", [(2, 8), (2, 4), (3, 6), (3, 8)])
def test_oracle_matches_reference_core_live(ndim, order):
    rng = np.random.default_rng(7 + order)
    nt = 120
    if ndim == 2:
        shape = (24, 30)
        P = oracle.A2DProblem(order, shape, 8.0, 8e-4, nt, abcn=9, B=1)
        src = np.array([[[11, 2]]])
        rec = np.array([[[i, 1] for i in range(0, 30, 4)]])
    else:
        shape = (12, 10, 14)
        P = oracle.A3DProblem(order, shape, 8.0, 8e-4, nt, abcn=7, B=1)
        src = np.array([[[6, 4, 2]]])
        rec = np.array([[[i, 5, 1] for i in range(0, 14, 3)]])
    vp = (2000 + 400 * rng.random(shape)).astype(np.float32)
    t = np.arange(nt) * 8e-4 - 0.06
    a = (np.pi * 20.0 * t) ** 2
    wav = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)[None, None]
    res = rng.standard_normal((1, rec.shape[1], nt)).astype(np.float32)
    vpp, s, r = P.pad_model(vp), P.shift(src), P.shift(rec)
    ref = refcore.run(P, vpp, wav, s, r, res)
    o = P.gradient(0, vpp, wav, s, r, res)
    bar = 1e-6 if ndim == 3 else 5e-5
    assert rel_l2(o["rec"], ref["rec"]) <= bar
    assert rel_l2(o["grad_vp"], ref["grad_vp"]) <= bar
    assert rel_l2(o["grad_wavelet"], ref["grad_wavelet"]) <= bar


# ---------------------------------------------------------------- elastic 2-D
E2D_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore_e2d_o*.npz")))
E2D_BS_CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(str(GOLDEN / "refcore_e2d_bs_*.npz")))


def e2d_problem(z, bs=False):
    B, nt = int(z["B"]), int(z["nt"])
    P = oracle.E2DProblem(int(z["order"]), tuple(int(v) for v in z["shape"]), float(z["dh"]), float(z["dt"]), nt,
                          abcn=int(z["abcn"]), B=B, bs=bs)
    wav = np.broadcast_to(np.asarray(z["wavelet"], np.float32), (B, 1, nt)).copy()
    models = [P.pad_model(z[k]) for k in ("vp", "vs", "rho")]
    return P, models, wav, P.shift(z["sources"]), P.shift(z["receivers"])


def test_e2d_fixtures_present():
    assert len(E2D_CASES) >= 3 and len(E2D_BS_CASES) >= 3


@pytest.mark.parametrize("case", E2D_BS_CASES)
def test_e2d_boundary_saving_matches_reference_core(case):
    """The reference's boundary-saving backward (backward_bs_core) against the
    oracle's boundary-saving loop (it = nt-1 .. 1), which images against the
    stored forward velocities: the reference rebuilds them by reverse steps
    and saved strips, so they differ by that reconstruction's rounding
    (measured <= 2.8e-5); records are the forward's (1e-6)."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, (vp, vs, rho), wav, s, r = e2d_problem(z, bs=True)
    o = P.gradient(vp, vs, rho, wav, s, r, z["residual"])
    assert rel_l2(o["rec"], z["rec"]) <= 1e-6
    for k in ("grad_vp", "grad_vs", "grad_rho"):
        assert rel_l2(P.crop(o[k]), z[k]) <= 1e-4, k


@pytest.mark.parametrize("case", E2D_CASES)
def test_e2d_oracle_matches_reference_core(case):
    """Elastic records agree to 1e-6 and the vp / vs / rho gradients to 5e-5
    (order 2 bit-exact; the higher fixed orders differ from the reference
    core in the rounding of a few constant products)."""
    z = np.load(GOLDEN / f"{case}.npz")
    P, (vp, vs, rho), wav, s, r = e2d_problem(z)
    o = P.gradient(vp, vs, rho, wav, s, r, z["residual"])
    assert rel_l2(o["rec"], z["rec"]) <= 1e-6
    for k in ("grad_vp", "grad_vs", "grad_rho"):
        assert rel_l2(P.crop(o[k]), z[k]) <= 5e-5, k


@pytest.mark.parametrize("model", [0, 1, 2], ids=["vp", "vs", "rho"])
def test_e2d_oracle_gradient_directional_fd(model):
    """<grad, dm> = dJ/de (J = <record, residual>) by central differences,
    for a smooth perturbation kept away from the sources and receivers."""
    z = np.load(GOLDEN / "refcore_e2d_o8.npz")
    P, models, wav, s, r = e2d_problem(z)
    res = z["residual"]
    o = P.gradient(*models, wav, s, r, res)
    shape = tuple(int(v) for v in z["shape"])
    rng = np.random.default_rng(11 + model)
    from scipy.ndimage import gaussian_filter
    d = gaussian_filter(rng.standard_normal(shape), 3.0)
    d[:8] = 0.0
    d[-4:] = 0.0
    d[:, :4] = 0.0
    d[:, -4:] = 0.0
    d = (d / np.abs(d).max()).astype(np.float32)
    base = [z["vp"], z["vs"], z["rho"]][model]
    eps = 1e-3 * float(np.abs(base).mean())

    def J(sign):
        ms = [np.asarray(z[k], np.float32).copy() for k in ("vp", "vs", "rho")]
        ms[model] = (ms[model] + sign * eps * d).astype(np.float32)
        rec = P.forward(*[P.pad_model(m) for m in ms], wav, s, r)
        return float(np.sum(rec.astype(np.float64) * res))

    fd = (J(1.0) - J(-1.0)) / (2 * eps)
    key = ["grad_vp", "grad_vs", "grad_rho"][model]
    an = float(np.sum(P.crop(o[key]).sum(0).astype(np.float64) * d))
    assert abs(fd - an) <= 2e-3 * abs(an), (fd, an)
