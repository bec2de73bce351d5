"""Write golden vectors from the reference's compiled acoustic path.

Run in the build container only, after ``oracle/build_ref.sh`` (the reference
tree is not present on GPU machines)::

    cd /tmp && python /root/repo/tests/golden/gen_refcore.py           # every fixture
    cd /tmp && python /root/repo/tests/golden/gen_refcore.py e2d_bs    # the elastic boundary-saving ones
    cd /tmp && python /root/repo/tests/golden/gen_refcore.py a2d_bs    # the 2-D acoustic boundary-saving ones
    cd /tmp && python /root/repo/tests/golden/gen_refcore.py a3d_bs    # the 3-D acoustic boundary-saving ones
    cd /tmp && python /root/repo/tests/golden/gen_refcore.py illum     # the illumination ones

The reference's CPU host core (its own CUDA sources compiled as C++, see
``oracle/build_ref.sh``) runs each case in full-storage mode through
``oracle/ref_harness.cpp``: the forward record, then the adjoint driven by a
seeded residual, giving the model gradient on the runtime grid and the
wavelet gradient.  The elastic cases are also run with boundary saving
(``refcore_e2d_bs_*``: the reference's backward_bs_core rebuilding the
forward state from its saved strips), and the acoustic cases too
(``refcore{2,3}d_bs_*``).  ``refcore{2,3}d_illum[_bs]_*`` add the source and
receiver illumination (compute_illumination) of one case per dimension.  Problem assembly (padding, coordinate shift, CPML
profiles, FD tables) is the oracle's, whose profiles and tables the table
tests pin bit-exactly against the reference's own builders.  Only inputs and
outputs are stored, as ``.npz`` arrays.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent
ROOT = OUT.parents[1]


def ricker(t, f):
    a = (np.pi * f * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def smooth_model(rng, shape, lo, hi):
    from scipy.ndimage import gaussian_filter
    v = gaussian_filter(rng.standard_normal(shape), sigma=2.0)
    v = (v - v.min()) / max(v.max() - v.min(), 1e-12)
    z = np.linspace(0.0, 1.0, shape[0]).reshape((-1,) + (1,) * (len(shape) - 1))
    return (lo + (hi - lo) * (0.6 * z + 0.4 * v)).astype(np.float32)


CASES_2D = [
    # name, order, shape, nsrc, B, nt
    ("o2", 2, (32, 44), 1, 2, 300),
    ("o4", 4, (32, 44), 1, 2, 300),
    ("o6", 6, (32, 44), 2, 2, 300),
    ("o8", 8, (32, 44), 1, 3, 300),
]
CASES_E2D = [
    # name, order, shape, B, nt
    ("o2", 2, (36, 48), 2, 300),
    ("o4", 4, (36, 48), 2, 300),
    ("o8", 8, (36, 48), 2, 300),
]
CASES_3D = [
    ("o2", 2, (14, 12, 16), 1, 2, 150),
    ("o4", 4, (14, 12, 16), 1, 2, 150),
    ("o8", 8, (14, 12, 16), 1, 2, 150),
]


def main(only=None):
    if os.getcwd().startswith("/root/reference"):
        raise SystemExit("run from outside /root/reference")
    sys.path.insert(0, str(ROOT))
    from oracle import oracle, refcore
    if not refcore.available():
        raise SystemExit("oracle/_ref is not built: run oracle/build_ref.sh first")
    dh, dt, abcn, f0 = 10.0, 1e-3, 10, 15.0
    acoustic = {None: ((2, CASES_2D), (3, CASES_3D)), "a2d_bs": ((2, CASES_2D),),
                "a3d_bs": ((3, CASES_3D),), "illum": ((2, CASES_2D), (3, CASES_3D))}.get(only, ())
    illum_cases = {(2, "o8"), (3, "o4")}
    for ndim, cases in acoustic:
        for name, order, shape, nsrc, B, nt in cases:
            rng = np.random.default_rng(1000 * ndim + order)
            ab = abcn if ndim == 2 else 8
            if ndim == 2:
                P = oracle.A2DProblem(order, shape, dh, dt, nt, abcn=ab, B=B)
            else:
                P = oracle.A3DProblem(order, shape, dh, dt, nt, abcn=ab, B=B)
            vp = smooth_model(rng, shape, 1800.0, 2600.0)
            wav = ricker(np.arange(nt) * dt - 1.2 / f0, f0)
            nx = shape[-1]
            src = np.zeros((B, nsrc, ndim), np.int64)
            for b in range(B):
                for s in range(nsrc):
                    src[b, s, 0] = int(rng.integers(2, nx - 2))
                    if ndim == 3:
                        src[b, s, 1] = int(rng.integers(2, shape[1] - 2))
                    src[b, s, -1] = int(rng.integers(1, 4))
            nrec = 7
            rec = np.zeros((B, nrec, ndim), np.int64)
            rec[..., 0] = np.linspace(0, nx - 1, nrec).astype(np.int64)
            if ndim == 3:
                rec[..., 1] = shape[1] // 2
            rec[..., -1] = 2
            wavB = np.broadcast_to(wav, (B, nsrc, nt)).copy()
            res = rng.standard_normal((B, nrec, nt)).astype(np.float32)
            if only in (None, "illum") and (ndim, name) in illum_cases:
                for bs in (False, True):
                    out = refcore.run_illum(P, P.pad_model(vp), wavB, P.shift(src), P.shift(rec), res, bs=bs)
                    path = OUT / f"refcore{ndim}d_illum{'_bs' if bs else ''}_{name}.npz"
                    np.savez_compressed(path, order=order, shape=np.array(shape), dh=dh, dt=dt, nt=nt, abcn=ab,
                                        B=B, vp=vp, wavelet=wav, sources=src, receivers=rec, residual=res,
                                        rec=out["rec"], grad_vp=out["grad_vp"], grad_wavelet=out["grad_wavelet"],
                                        illum_s=out["illum_s"], illum_r=out["illum_r"])
                    print(path.name, os.path.getsize(path))
            if only == "illum":
                continue
            for bs in ((False, True) if only is None else (True,)):
                out = refcore.run(P, P.pad_model(vp), wavB, P.shift(src), P.shift(rec), res, bs=bs)
                path = OUT / (f"refcore{ndim}d_bs_{name}.npz" if bs else f"refcore{ndim}d_{name}.npz")
                np.savez_compressed(path, order=order, shape=np.array(shape), dh=dh, dt=dt, nt=nt, abcn=ab, B=B,
                                    vp=vp, wavelet=wav, sources=src, receivers=rec, residual=res,
                                    rec=out["rec"], grad_vp=out["grad_vp"], grad_wavelet=out["grad_wavelet"])
            print(path.name, os.path.getsize(path))
    # elastic 2-D: vz sources, vx + vz receivers, layered vp with vs = vp / 1.8, Gardner rho
    modes = {None: (False, True), "e2d_bs": (True,)}.get(only, ())
    dh, dt, abcn, f0 = 5.0, 5e-4, 10, 15.0
    for name, order, shape, B, nt in CASES_E2D:
        rng = np.random.default_rng(3000 + order)
        P = oracle.E2DProblem(order, shape, dh, dt, nt, abcn=abcn, B=B)
        vp = smooth_model(rng, shape, 2500.0, 3800.0)
        vs = (vp / 1.8 * (1.0 + 0.05 * rng.uniform(-1, 1, shape))).astype(np.float32)
        rho = (310.0 * vp ** 0.25).astype(np.float32)
        wav = ricker(np.arange(nt) * dt - 1.2 / f0, f0)
        nx = shape[-1]
        src = np.stack([rng.integers(3, nx - 3, B), rng.integers(1, 4, B)], -1)[:, None, :]
        nrec = 9
        rec = np.zeros((B, nrec, 2), np.int64)
        rec[..., 0] = np.linspace(0, nx - 1, nrec).astype(np.int64)
        rec[..., 1] = 2
        wavB = np.broadcast_to(wav, (B, 1, nt)).copy()
        res = rng.standard_normal((2, B, nrec, nt)).astype(np.float32)
        for bs in modes:
            out = refcore.run_elastic((P.nz, P.nx), P.M, nt, abcn, dt, (dh, dh), P.prof, P.pad_model(vp),
                                      P.pad_model(vs), P.pad_model(rho), P.grad, wavB, P.shift(src), P.shift(rec),
                                      res, bs=bs)
            path = OUT / (f"refcore_e2d_bs_{name}.npz" if bs else f"refcore_e2d_{name}.npz")
            np.savez_compressed(path, order=order, shape=np.array(shape), dh=dh, dt=dt, nt=nt, abcn=abcn, B=B,
                                vp=vp, vs=vs, rho=rho, wavelet=wav, sources=src, receivers=rec, residual=res,
                                rec=out["rec"], grad_vp=P.crop(out["grad_vp"]), grad_vs=P.crop(out["grad_vs"]),
                                grad_rho=P.crop(out["grad_rho"]))
            print(path.name, os.path.getsize(path))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
