"""Write the acoustic 2-D golden vectors from the reference implementation.

Run in the build container only (the reference is not present on GPU
machines):  ``cd /tmp && python /root/repo/tests/golden/gen_acoustic2d.py``.

The reference's eager PyTorch path (``impl='eager'``, CPU, ``Full()`` memory,
``EagerOptions(use_compile=False)``) computes each case; autograd gives the
gradients of ``0.5 * ||pred - obs||^2`` with respect to the starting model and
the wavelet.  The CPML profiles and FD tables come from the reference's own
builders (``equations/pml.py``, ``scalars.py``).  Nothing of the reference is
stored: only inputs and outputs, as ``.npz`` arrays.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

REF = "/root/reference/src"
OUT = Path(__file__).resolve().parent


def _import_reference():
    if os.getcwd().startswith("/root/reference"):
        raise SystemExit("run from outside /root/reference (its signal.py shadows the stdlib)")
    sys.dont_write_bytecode = True
    sys.path.insert(0, REF)
    import torch  # noqa: F401
    from sweep.equations import Acoustic
    from sweep.propagator.torch import PropTorch
    from sweep.propagator.options import EagerOptions, Full
    from sweep import scalars
    from sweep.equations import pml
    return Acoustic, PropTorch, EagerOptions, Full, scalars, pml


def ricker(t, f):
    r = (1 - 2 * (np.pi * f * t) ** 2) * np.exp(-(np.pi * f * t) ** 2)
    return r.astype(np.float32)


def layered_model(rng, shape, vmin, vmax, nlayers):
    nz, nx = shape
    z = np.arange(nz)[:, None].astype(np.float64)
    x = np.arange(nx)[None, :].astype(np.float64)
    depths = np.sort(rng.uniform(0.15 * nz, 0.9 * nz, nlayers))
    vel = np.sort(rng.uniform(vmin, vmax, nlayers + 1))
    vp = np.full(shape, vel[0])
    for k, d in enumerate(depths):
        amp = rng.uniform(0.5, 2.5)
        lam = rng.uniform(0.5, 1.5) * nx
        phase = rng.uniform(0, 2 * np.pi)
        iface = d + amp * np.sin(2 * np.pi * x / lam + phase)
        vp = np.where(z >= iface, vel[k + 1], vp)
    noise = rng.standard_normal(shape)
    k = np.array([1, 4, 6, 4, 1], dtype=np.float64)
    k /= k.sum()
    for ax in (0, 1):
        noise = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, noise)
    noise /= np.abs(noise).max()
    return (vp * (1 + 0.05 * noise)).astype(np.float32)


CASES = {
    # the README example at the smallest configuration, in full
    "c1_order4": dict(order=4, shape=(96, 128), dh=10.0, dt=2e-3, nt=800, abcn=50, f=10.0,
                      delay=0.14, src=[[32, 48]], rec=[[[96, 48]]], model="two_layer"),
    # heterogeneous, order 8, two shots
    "het_order8": dict(order=8, shape=(40, 60), dh=8.0, dt=8e-4, nt=400, abcn=16, f=25.0,
                       delay=0.05, src=[[12, 3], [45, 4]],
                       rec=[[[4 + 9 * r, 2] for r in range(6)], [[7 + 8 * r, 3] for r in range(6)]],
                       model=("layered", 11, 1500.0, 4500.0, 4)),
    "het_order6": dict(order=6, shape=(36, 44), dh=5.0, dt=6e-4, nt=300, abcn=10, f=40.0,
                       delay=0.03, src=[[20, 5]], rec=[[[3 + 7 * r, 30] for r in range(6)]],
                       model=("layered", 12, 1800.0, 4000.0, 3)),
    "het_order2": dict(order=2, shape=(30, 40), dh=10.0, dt=1e-3, nt=300, abcn=8, f=20.0,
                       delay=0.06, src=[[10, 12]], rec=[[[30, 14], [5, 25], [36, 2]]],
                       model=("layered", 13, 1500.0, 4000.0, 2)),
    # runtime-order path (M = 5)
    "het_order10": dict(order=10, shape=(30, 36), dh=10.0, dt=1e-3, nt=200, abcn=8, f=20.0,
                        delay=0.06, src=[[15, 10]], rec=[[[2, 3], [33, 20], [17, 27]]],
                        model=("layered", 14, 1500.0, 4000.0, 2)),
}


def make_models(spec, shape):
    if spec == "two_layer":
        true = np.full(shape, 1500.0, dtype=np.float32)
        true[shape[0] // 2:, :] = 2500.0
        init = np.full(shape, 1500.0, dtype=np.float32)
        return true, init
    _, seed, vmin, vmax, nl = spec
    rng = np.random.default_rng(seed)
    true = layered_model(rng, shape, vmin, vmax, nl)
    # starting model: the true one smoothed hard
    k = np.ones(9) / 9.0
    init = true.astype(np.float64)
    for _ in range(3):
        for ax in (0, 1):
            init = np.apply_along_axis(lambda v: np.convolve(np.pad(v, 4, mode="edge"), k, "valid"),
                                       ax, init)
    return true, init.astype(np.float32)


def run_case(name, c, mods):
    Acoustic, PropTorch, EagerOptions, Full, _, _ = mods
    import torch
    torch.set_num_threads(8)
    shape = c["shape"]
    vp_true, vp_init = make_models(c["model"], shape)
    t = np.arange(c["nt"], dtype=np.float64) * c["dt"]
    wavelet = ricker(t - c["delay"], c["f"])
    src = np.array(c["src"], dtype=np.int64)
    rec = np.array(c["rec"], dtype=np.int64)
    solver = PropTorch(Acoustic(spatial_order=c["order"], device="cpu"), shape=shape, dh=c["dh"],
                       dt=c["dt"], device="cpu", impl="eager", abcn=c["abcn"], memory=Full(),
                       eager_options=EagerOptions(use_compile=False))
    with torch.no_grad():
        obs = solver(wavelet, src, rec, models=[torch.tensor(vp_true)])
    vp_t = torch.tensor(vp_init, requires_grad=True)
    w_t = torch.tensor(wavelet, requires_grad=True)
    pred = solver(w_t, src, rec, models=[vp_t])
    loss = 0.5 * (pred - obs).pow(2).sum()
    loss.backward()
    prof = [np.asarray(p.detach().cpu().numpy() if hasattr(p, "detach") else p, dtype=np.float32).ravel()
            for p in solver.equation.b]
    out = dict(order=c["order"], shape=np.array(shape), dh=c["dh"], dt=c["dt"], nt=c["nt"],
               abcn=c["abcn"], vp_true=vp_true, vp_init=vp_init, wavelet=wavelet, sources=src,
               receivers=rec, obs=obs.numpy(), pred=pred.detach().numpy(),
               grad_vp=vp_t.grad.numpy(), grad_wavelet=w_t.grad.numpy(), loss=float(loss),
               az=prof[0], bz=prof[1], dbz=prof[2], ax=prof[3], bx=prof[4], dbx=prof[5])
    np.savez_compressed(OUT / f"acoustic2d_{name}.npz", **out)
    print(name, "loss", float(loss), "obs max", float(obs.abs().max()), flush=True)


def write_tables(mods):
    _, _, _, _, scalars, pml = mods
    tabs = {}
    for M in range(1, 9):
        c0, c = scalars.second_derivative_coefficients_fp32(2 * M)
        tabs[f"lap_c0_{M}"] = np.float32(c0)
        tabs[f"lap_c_{M}"] = c
        tabs[f"grad_{M}"] = scalars.fd_coefficients(1, 2 * M).astype(np.float32)
        tabs[f"stag_{M}"] = scalars.staggered_grid_coes(M)
    # CPML profiles at the full runtime geometries (order 8, abcn 50): the 2-D
    # acoustic grid 384x1152 (dh 8, dt 8e-4), the elastic 512x1024 (dh 5, dt
    # 5e-4, staggered), the 3-D 256x512x512 (dh 10, dt 1e-3).
    M = 4
    geo = {
        "r_384x1152": ("r", (384, 1152), 8.0, 8e-4),
        "s_512x1024": ("s", (512, 1024), 5.0, 5e-4),
        "r_256x512x512": ("r", (256, 512, 512), 10.0, 1e-3),
        "r_96x128_o4": ("r", (96, 128), 10.0, 2e-3),
    }
    for key, (kind, shp, dh, dt) in geo.items():
        m = 2 if key.endswith("_o4") else M
        runtime = [s + 2 * 50 + 2 * m for s in shp]
        args = dict(pml_width=[50] * (2 * len(shp)), accuracy=2 * m, fd_pad=[m, m] * len(shp),
                    dt=dt, grid_spacing=[dh] * len(shp), max_vel=4500.0, dtype=np.float32,
                    pml_freq=25.0, shape=tuple(runtime))
        fn = pml.set_cpml_profiles_r if kind == "r" else pml.set_cpml_profiles_s
        for i, p in enumerate(fn(**args)):
            tabs[f"prof_{key}_{i}"] = np.asarray(p, dtype=np.float32).ravel()
    np.savez_compressed(OUT / "tables.npz", **tabs)
    print("tables written", len(tabs), flush=True)


def main():
    mods = _import_reference()
    write_tables(mods)
    only = sys.argv[1:]
    for name, c in CASES.items():
        if only and name not in only:
            continue
        run_case(name, c, mods)


if __name__ == "__main__":
    main()
