"""Pinning the 2-D acoustic oracle (CPU, no GPU needed).

The oracle (oracle/acoustic2d.c) restates the reference's compiled path.  It
is pinned by
  1. the reference's own records and gradients (tests/golden/, written by the
     reference's eager PyTorch path): within the reference's compiled-vs-eager
     tolerance (test/test_host_core.py:229, 1e-3);
  2. the same goldens to ~1e-6 once the oracle mirrors the eager path's outer
     edge (the only place the two reference paths differ: the compiled path
     never updates the outer M cells of the runtime grid, the eager path
     convolves them with zero padding) -- so interior, CPML, source and
     receiver arithmetic are the reference's;
  3. the dot-product test (the adjoint is the exact transpose of the forward
     in the wavelet), a directional finite-difference check of the model
     gradient, Full == boundary-saving agreement, and the analytic 2-D
     Green's function of a homogeneous medium.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, cosine, rel_l2
from oracle.oracle import A2DProblem, a2d_from_golden, lib

CASES = ["c1_order4", "het_order8", "het_order6", "het_order2", "het_order10"]


def _raw(canon):
    return np.asarray(canon)[..., 0].transpose(0, 2, 1)


@pytest.fixture(scope="module")
def runs():
    out = {}
    for name in CASES:
        z = np.load(GOLDEN / f"acoustic2d_{name}.npz")
        P, wav, sl, rl = a2d_from_golden(z)
        obs = P.forward(P.pad_model(z["vp_true"]), wav, sl, rl)
        vpi = P.pad_model(z["vp_init"])
        pred = P.forward(vpi, wav, sl, rl)
        g = {m: P.gradient(m, vpi, wav, sl, rl, pred - obs) for m in (0, 1)}
        out[name] = dict(z=z, P=P, obs=obs, pred=pred,
                         grad={m: P.crop(g[m]["grad_vp"]).sum(0) for m in g},
                         gw={m: g[m]["grad_wavelet"].sum((0, 1)) for m in g})
    return out


@pytest.mark.parametrize("name", CASES)
def test_oracle_vs_reference_golden(runs, name):
    r = runs[name]
    z = r["z"]
    assert rel_l2(r["obs"], _raw(z["obs"])) <= 1e-3
    assert rel_l2(r["pred"], _raw(z["pred"])) <= 1e-3
    for m in (0, 1):
        assert cosine(r["grad"][m], z["grad_vp"]) >= 0.9999
        assert rel_l2(r["grad"][m], z["grad_vp"]) <= 1e-3
        assert rel_l2(r["gw"][m], z["grad_wavelet"]) <= 1e-3


@pytest.mark.parametrize("name", CASES)
def test_oracle_with_eager_edge_matches_golden_tightly(name):
    z = np.load(GOLDEN / f"acoustic2d_{name}.npz")
    P, wav, sl, rl = a2d_from_golden(z)
    lib().orc_set_edge_mode(1)
    try:
        obs = P.forward(P.pad_model(z["vp_true"]), wav, sl, rl)
        pred = P.forward(P.pad_model(z["vp_init"]), wav, sl, rl)
    finally:
        lib().orc_set_edge_mode(0)
    assert rel_l2(obs, _raw(z["obs"])) <= 5e-5
    assert rel_l2(pred, _raw(z["pred"])) <= 5e-5


def test_edge_mode_is_what_differs(runs):
    """Without the eager edge the c1 record sits ~5e-4 from eager, with it
    ~2e-5: the residual is the outer-edge treatment, nothing else."""
    r = runs["c1_order4"]
    z = r["z"]
    P, wav, sl, rl = a2d_from_golden(z)
    lib().orc_set_edge_mode(1)
    try:
        obs_e = P.forward(P.pad_model(z["vp_true"]), wav, sl, rl)
    finally:
        lib().orc_set_edge_mode(0)
    far = rel_l2(r["obs"], _raw(z["obs"]))
    near = rel_l2(obs_e, _raw(z["obs"]))
    assert far > 1e-4 and near < far / 10


@pytest.mark.parametrize("name", CASES)
def test_full_and_boundary_saving_agree(runs, name):
    r = runs[name]
    assert rel_l2(r["grad"][1], r["grad"][0]) <= 1e-4
    assert np.array_equal(r["gw"][1], r["gw"][0])


@pytest.mark.parametrize("order", [2, 4, 8, 10])
def test_dot_product_adjoint(order):
    """<F w, d> == <w, F^T d>: the wavelet gradient is the exact transpose of
    the forward (linear in the wavelet), CPML and point operators included."""
    rng = np.random.default_rng(order)
    shape, nt = (30, 34), 150
    P = A2DProblem(order, shape, 10.0, 1e-3, nt, abcn=8, B=1)
    vp = P.pad_model(rng.uniform(1500, 3500, shape).astype(np.float32))
    w = rng.standard_normal((1, 2, nt)).astype(np.float32)
    sl = P.shift([[[3, 2], [25, 28]]])
    rl = P.shift([[[1, 1], [14, 9], [33, 29], [20, 0]]])
    d = rng.standard_normal((1, 4, nt)).astype(np.float32)
    Fw = P.forward(vp, w, sl, rl)
    for mode in (0, 1):
        Ftd = P.gradient(mode, vp, w, sl, rl, d)["grad_wavelet"]
        lhs = float(np.dot(Fw.ravel().astype(np.float64), d.ravel()))
        rhs = float(np.dot(w.ravel().astype(np.float64), Ftd.ravel()))
        assert abs(lhs - rhs) <= 2e-5 * max(abs(lhs), 1e-30), (mode, lhs, rhs)


def test_model_gradient_directional_fd():
    """Central finite difference of 0.5||F(m) - d||^2 (accumulated in float64)
    along a smooth direction vs <grad, dm>."""
    z = np.load(GOLDEN / "acoustic2d_het_order8.npz")
    P, wav, sl, rl = a2d_from_golden(z)
    vp0 = np.asarray(z["vp_init"], dtype=np.float64)
    obs = P.forward(P.pad_model(z["vp_true"]), wav, sl, rl)
    rng = np.random.default_rng(3)
    dm = rng.standard_normal(vp0.shape)
    k = np.ones(7) / 7
    for ax in (0, 1):
        dm = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, dm)
    # zero within 6 cells of the physical edge: the edge values are replicated
    # into the PML pad, whose gradient the propagator crops by design
    # (utils/torch.py:19-41), so a perturbation there is not what it images
    taper = np.zeros_like(dm)
    taper[6:-6, 6:-6] = 1.0
    dm *= taper
    dm *= 5.0 / np.abs(dm).max()           # up to 5 m/s, ~0.2% of vp

    def loss(m):
        pred = P.forward(P.pad_model(m.astype(np.float32)), wav, sl, rl)
        return 0.5 * float(np.sum((pred.astype(np.float64) - obs) ** 2)), pred

    _, pred0 = loss(vp0)
    g = P.crop(P.gradient(1, P.pad_model(vp0.astype(np.float32)), wav, sl, rl, pred0 - obs)["grad_vp"]).sum(0)
    lp, _ = loss(vp0 + dm)
    lm, _ = loss(vp0 - dm)
    fd = (lp - lm) / 2.0
    an = float(np.sum(g.astype(np.float64) * dm))
    assert abs(fd - an) <= 1e-2 * abs(fd), (fd, an)


def _ricker(t, f0):
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def _green(times, V, OFF, f0, n=4001):
    out = np.zeros(len(times))
    for k, t in enumerate(times):
        if V * t > OFF:
            u = np.linspace(0.0, np.arccosh(V * t / OFF), n)
            y = _ricker(t - OFF * np.cosh(u) / V, f0)
            out[k] = (y.sum() - 0.5 * (y[0] + y[-1])) * (u[1] - u[0]) / (2 * np.pi * V ** 2)
    return out


@pytest.mark.parametrize("order", [4, 8])
def test_record_converges_to_analytic_green_function(order):
    """Homogeneous 2000 m/s, Ricker 10 Hz delayed 1.2/f, offset 300 m: the
    geometry of the reference's test/test_laplacian_zero_sum.py:83-122 on a
    2.5 m grid (that test uses 1.25 m on a GPU and bounds the error by
    1.5e-4; at 2.5 m the oracle measures 1.6e-4 / 1.7e-4 for orders 4 / 8)."""
    V, F0, T, L, OFF = 2000.0, 10.0, 0.45, 1600.0, 300.0
    dh = 2.5
    dt = 1e-4 * dh
    nt, n = int(round(T / dt)), int(round(L / dh)) + 1
    s = (n - 1) // 2
    P = A2DProblem(order, (n, n), dh, dt, nt, abcn=30, B=1)
    w = np.ascontiguousarray(_ricker(np.arange(nt) * dt, F0).astype(np.float32)[None, None, :])
    d = P.forward(P.pad_model(np.full((n, n), V, np.float32)), w, P.shift([[[s, s]]]),
                  P.shift([[[s + int(round(OFF / dh)), s]]]))
    trace = d[0, 0].astype(np.float64) * dt ** 2 / dh ** 2
    ref = _green((np.arange(nt) + 1) * dt, V, OFF, F0)
    assert rel_l2(trace, ref) < 2.5e-4


def test_fma_contraction_sensitivity_of_c1():
    """The c1 case (two-layer model, order 4, receiver on the interface) is
    rounding-sensitive: the oracle's own source built with FMA contraction
    moves its record by > 1e-5 (4.1e-5 measured), while the other cases move
    by ~1e-6.  The reference's CUDA core is built with --use_fast_math
    (backend/c/jit.py:1065), so contraction is part of the reference too; the
    GPU parity test takes this measured floor as its bar for that record."""
    gaps = {}
    for name in ("c1_order4", "het_order8"):
        z = np.load(GOLDEN / f"acoustic2d_{name}.npz")
        Pi, wav, sl, rl = a2d_from_golden(z, "ieee")
        Pf, _, _, _ = a2d_from_golden(z, "fma")
        a = Pi.forward(Pi.pad_model(z["vp_true"]), wav, sl, rl)
        b = Pf.forward(Pf.pad_model(z["vp_true"]), wav, sl, rl)
        gaps[name] = rel_l2(b, a)
    assert gaps["c1_order4"] > 1e-5
    assert gaps["het_order8"] < 5e-6


def test_bf16_strip_rounding_is_round_to_nearest_even():
    """The oracle's strip rounding against torch's float32 -> bfloat16
    conversion (round to nearest, ties to even): ties in both directions,
    subnormals, overflow to inf, signed zeros, inf and a NaN."""
    import torch
    from oracle.oracle import lib
    L = lib()
    rng = np.random.default_rng(7)
    xs = np.concatenate([
        rng.standard_normal(2000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 2000),
        np.array([1.0, 1.00390625, 1.005859375, 1.0078125, 1.01171875, -1.00390625, 1e-40, -1e-41, 3.39e38,
                  3.4e38, np.inf, -np.inf, 0.0, -0.0], dtype=np.float32)]).astype(np.float32)
    ref = torch.tensor(xs).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([L.orc_bf16_round(float(x)) for x in xs], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(L.orc_bf16_round(float("nan")))
    assert L.orc_bf16_round(1.00390625) == 1.0 and L.orc_bf16_round(1.01171875) == 1.015625   # ties to even
