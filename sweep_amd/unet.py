"""A 2-D U-Net on the HIP kernels of libswp_core (bf16 matrix-core
convolutions, float32 accumulation): the learned preconditioner of FWI.

Architecture (``levels`` resolutions, ``base`` channels at the finest,
doubling per level; ReLU after every 3x3 convolution):

* encoder level l: conv3x3 -> conv3x3 (output e_l), then a 2x2 average pool
  into level l + 1 (not below the last level);
* decoder level l (from levels - 2 up to 0): nearest 2x upsample of the level
  below, concatenated with e_l, conv3x3 -> conv3x3;
* a 1x1 convolution to one output channel (float32).

Activations and weights are bfloat16, NHWC; the input has 8 channels (the
``nsrc`` input planes, scaled, then zeros) and the output layer 8 (channel 0
is the output), so every convolution's channel count is a multiple of 8.
The concatenations are never copied: e_l is written straight into the
channel slice of its decoder's input buffer.  Master weights are float32 and
trained with Adam; the kernels read bf16 copies refreshed by each step.

Everything runs on the GPU; there is no host path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _core

BF16 = torch.bfloat16
IN_CH = 8


def _p(t):
    return None if t is None else t.data_ptr()


def _call(name, *args):
    L = _core.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    _core.on_stream(dev, getattr(L, name), *args)


@dataclass
class Layer:
    name: str
    cin: int
    cout: int
    ks: int
    w: torch.Tensor          # float32 [cout, ks*ks, cin]
    b: torch.Tensor          # float32 [cout]
    wbf: torch.Tensor        # bf16 copy of w
    gw: torch.Tensor = None
    gb: torch.Tensor = None
    mw: torch.Tensor = None
    vw: torch.Tensor = None
    mb: torch.Tensor = None
    vb: torch.Tensor = None
    wt: torch.Tensor = None  # flipped bf16 weights [cin, ks*ks, cout] of the data gradient


class UNet:

    def __init__(self, device, levels: int = 4, base: int = 32, nsrc: int = 3, seed: int = 0):
        if nsrc > 4:
            raise ValueError("at most 4 input planes")
        self.device = torch.device(device)
        self.levels, self.base, self.nsrc = levels, base, nsrc
        self.C = [base * 2 ** l for l in range(levels)]
        rng = np.random.default_rng(seed)
        specs = []
        for l in range(levels):
            specs.append((f"e{l}a", IN_CH if l == 0 else self.C[l - 1], self.C[l], 3))
            specs.append((f"e{l}b", self.C[l], self.C[l], 3))
        for l in range(levels - 2, -1, -1):
            specs.append((f"d{l}a", self.C[l + 1] + self.C[l], self.C[l], 3))
            specs.append((f"d{l}b", self.C[l], self.C[l], 3))
        specs.append(("out", self.C[0], IN_CH, 1))
        self.layers = {}
        for name, cin, cout, ks in specs:
            real_in = nsrc if name == "e0a" else cin
            w = np.zeros((cout, ks * ks, cin), np.float32)
            std = np.sqrt(2.0 / (real_in * ks * ks)) if name != "out" else 0.1 * np.sqrt(1.0 / cin)
            real_out = 1 if name == "out" else cout
            w[:real_out, :, :real_in] = rng.standard_normal((real_out, ks * ks, real_in)) * std
            wt = torch.from_numpy(w).to(self.device)
            lay = Layer(name, cin, cout, ks, wt, torch.zeros(cout, device=self.device),
                        torch.empty(wt.shape, dtype=BF16, device=self.device))
            for k in ("gw", "mw", "vw"):
                setattr(lay, k, torch.zeros_like(wt))
            for k in ("gb", "mb", "vb"):
                setattr(lay, k, torch.zeros_like(lay.b))
            lay.wt = torch.empty((cin, ks * ks, cout), dtype=BF16, device=self.device)
            self.layers[name] = lay
        self.refresh()
        self.step_count = 0
        self._acts = None
        self._scratch = None

    # -- parameters --------------------------------------------------------
    def refresh(self):
        """bf16 copies of the master weights."""
        for lay in self.layers.values():
            _call("swp_nn_to_bf16", _p(lay.w), _p(lay.wbf), lay.w.numel())

    def parameters(self):
        for lay in self.layers.values():
            yield lay.w
            yield lay.b

    def state(self):
        return {n: (l.w.detach().cpu().clone(), l.b.detach().cpu().clone()) for n, l in self.layers.items()}

    def save(self, path):
        """Master weights, biases and Adam moments (torch.save of plain tensors)."""
        out = {"step_count": torch.tensor(self.step_count)}
        for n, l in self.layers.items():
            for k in ("w", "b", "mw", "vw", "mb", "vb"):
                out[f"{n}.{k}"] = getattr(l, k).detach().cpu()
        torch.save(out, path)

    def load(self, path):
        """Restore what save() wrote (same architecture); refreshes the bf16 copies."""
        st = torch.load(path, map_location="cpu", weights_only=True)
        for n, l in self.layers.items():
            for k in ("w", "b", "mw", "vw", "mb", "vb"):
                t = st[f"{n}.{k}"]
                if tuple(t.shape) != tuple(getattr(l, k).shape):
                    raise ValueError(f"{n}.{k}: saved {tuple(t.shape)}, network {tuple(getattr(l, k).shape)}")
                getattr(l, k).copy_(t.to(self.device))
        self.step_count = int(st["step_count"])
        self.refresh()

    # -- building blocks ---------------------------------------------------
    def _conv(self, lay, x, xs, N, H, W, out, os, relu=True, f32=False):
        _call("swp_nn_conv_fwd", _p(x), N, H, W, lay.cin, xs, _p(lay.wbf), lay.cout, lay.ks, _p(lay.b),
              1 if relu else 0, _p(out), os, 1 if f32 else 0)

    def _empty(self, N, H, W, C):
        return torch.empty((N, H, W, C), dtype=BF16, device=self.device)

    def input(self, planes, scales):
        """(N, H, W, 8) bf16 network input from nsrc (N, H, W) float planes."""
        if len(planes) != self.nsrc or len(scales) != self.nsrc:
            raise ValueError(f"need {self.nsrc} planes and scales")
        planes = [p.to(self.device, torch.float32).contiguous() for p in planes]
        N, H, W = planes[0].shape
        x = self._empty(N, H, W, IN_CH)
        ptrs = (ctypes.c_void_p * self.nsrc)(*[p.data_ptr() for p in planes])
        sc = (ctypes.c_float * self.nsrc)(*[float(s) for s in scales])
        L = _core.load()
        _core.on_stream(self.device, L.swp_nn_pack_input, ctypes.cast(ptrs, ctypes.c_void_p),
                        ctypes.cast(sc, ctypes.c_void_p), self.nsrc, N * H * W, _p(x), IN_CH)
        return x, planes

    # -- forward -----------------------------------------------------------
    def forward(self, x):
        """x: (N, H, W, 8) bf16; returns the (N, H, W, 8) float32 output
        (channel 0) and keeps the activations for backward()."""
        N, H, W, _ = x.shape
        f = 2 ** (self.levels - 1)
        if H % f or W % f:
            raise ValueError(f"H and W must be multiples of {f} for {self.levels} levels")
        Ls, C = self.levels, self.C
        acts = {"x": x, "N": N, "H": H, "W": W}
        inp, ins = x, IN_CH
        for l in range(Ls):
            Hl, Wl = H >> l, W >> l
            a = self._empty(N, Hl, Wl, C[l])
            self._conv(self.layers[f"e{l}a"], inp, ins, N, Hl, Wl, a, C[l])
            if l < Ls - 1:
                cat = self._empty(N, Hl, Wl, C[l + 1] + C[l])
                e, es = cat[..., C[l + 1]:], C[l + 1] + C[l]
                acts[f"cat{l}"] = cat
            else:
                e = self._empty(N, Hl, Wl, C[l])
                es = C[l]
            self._conv(self.layers[f"e{l}b"], a, C[l], N, Hl, Wl, e, es)
            acts[f"in{l}"], acts[f"ins{l}"], acts[f"a{l}"], acts[f"e{l}"], acts[f"es{l}"] = inp, ins, a, e, es
            if l < Ls - 1:
                p = self._empty(N, Hl // 2, Wl // 2, C[l])
                _call("swp_nn_pool2_fwd", _p(e), es, N, Hl, Wl, C[l], _p(p), C[l])
                inp, ins = p, C[l]
        prev, prevs = acts[f"e{Ls - 1}"], C[Ls - 1]
        for l in range(Ls - 2, -1, -1):
            Hl, Wl = H >> l, W >> l
            cat = acts[f"cat{l}"]
            cs = C[l + 1] + C[l]
            _call("swp_nn_up2_fwd", _p(prev), prevs, N, Hl // 2, Wl // 2, C[l + 1], _p(cat), cs)
            da = self._empty(N, Hl, Wl, C[l])
            self._conv(self.layers[f"d{l}a"], cat, cs, N, Hl, Wl, da, C[l])
            db = self._empty(N, Hl, Wl, C[l])
            self._conv(self.layers[f"d{l}b"], da, C[l], N, Hl, Wl, db, C[l])
            acts[f"da{l}"], acts[f"db{l}"] = da, db
            prev, prevs = db, C[l]
        out = torch.empty((N, H, W, IN_CH), dtype=torch.float32, device=self.device)
        self._conv(self.layers["out"], prev, C[0], N, H, W, out, IN_CH, relu=False, f32=True)
        self._acts = acts
        return out

    # -- backward ----------------------------------------------------------
    def _wgrad(self, lay, x, xs, d, ds, N, H, W):
        if self._scratch is None or self._scratch.numel() < self._need(lay, N, H, W):
            self._scratch = torch.empty(self._need(lay, N, H, W), dtype=torch.float32, device=self.device)
        _call("swp_nn_conv_wgrad", _p(x), N, H, W, lay.cin, xs, _p(d), lay.cout, ds, lay.ks, _p(lay.gw), _p(lay.gb),
              _p(self._scratch))

    def _need(self, lay, N, H, W):
        return int(_core.load().swp_nn_conv_wgrad_scratch(N, H, W, lay.cin, lay.cout, lay.ks))

    def _dgrad(self, lay, d, ds, N, H, W, out, os):
        _call("swp_nn_conv_flip", _p(lay.wbf), lay.cout, lay.cin, lay.ks, _p(lay.wt))
        _call("swp_nn_conv_fwd", _p(d), N, H, W, lay.cout, ds, _p(lay.wt), lay.cin, lay.ks, None, 0, _p(out), os, 0)

    def backward(self, dout):
        """dout: (N, H, W, 8) bf16 gradient of the output (channel 0); fills
        every layer's gw / gb."""
        a = self._acts
        if a is None:
            raise RuntimeError("backward() needs a forward() first")
        N, H, W = a["N"], a["H"], a["W"]
        Ls, C = self.levels, self.C
        lay = self.layers["out"]
        self._wgrad(lay, a["db0"] if Ls > 1 else a[f"e{Ls - 1}"], C[0], dout, IN_CH, N, H, W)
        d = self._empty(N, H, W, C[0])
        self._dgrad(lay, dout, IN_CH, N, H, W, d, C[0])
        for l in range(0, Ls - 1):
            Hl, Wl = H >> l, W >> l
            cs = C[l + 1] + C[l]
            _call("swp_nn_relu_bwd", _p(d), C[l], _p(a[f"db{l}"]), C[l], N * Hl * Wl, C[l])
            lb = self.layers[f"d{l}b"]
            self._wgrad(lb, a[f"da{l}"], C[l], d, C[l], N, Hl, Wl)
            dda = self._empty(N, Hl, Wl, C[l])
            self._dgrad(lb, d, C[l], N, Hl, Wl, dda, C[l])
            _call("swp_nn_relu_bwd", _p(dda), C[l], _p(a[f"da{l}"]), C[l], N * Hl * Wl, C[l])
            la = self.layers[f"d{l}a"]
            self._wgrad(la, a[f"cat{l}"], cs, dda, C[l], N, Hl, Wl)
            dcat = self._empty(N, Hl, Wl, cs)
            self._dgrad(la, dda, C[l], N, Hl, Wl, dcat, cs)
            a[f"dcat{l}"] = dcat
            d = self._empty(N, Hl // 2, Wl // 2, C[l + 1])
            _call("swp_nn_up2_bwd", _p(dcat), cs, N, Hl // 2, Wl // 2, C[l + 1], _p(d), C[l + 1])
        # d: gradient of the deepest encoder output
        for l in range(Ls - 1, -1, -1):
            Hl, Wl = H >> l, W >> l
            if l == Ls - 1:
                de, des = d, C[l]
            else:
                de, des = a[f"dcat{l}"][..., C[l + 1]:], C[l + 1] + C[l]
                _call("swp_nn_pool2_bwd_acc", _p(dp), C[l], N, Hl, Wl, C[l], _p(de), des)
            _call("swp_nn_relu_bwd", _p(de), des, _p(a[f"e{l}"]), a[f"es{l}"], N * Hl * Wl, C[l])
            lb = self.layers[f"e{l}b"]
            self._wgrad(lb, a[f"a{l}"], C[l], de, des, N, Hl, Wl)
            dav = self._empty(N, Hl, Wl, C[l])
            self._dgrad(lb, de, des, N, Hl, Wl, dav, C[l])
            _call("swp_nn_relu_bwd", _p(dav), C[l], _p(a[f"a{l}"]), C[l], N * Hl * Wl, C[l])
            la = self.layers[f"e{l}a"]
            self._wgrad(la, a[f"in{l}"], a[f"ins{l}"], dav, C[l], N, Hl, Wl)
            if l > 0:
                dp = self._empty(N, Hl, Wl, C[l - 1])
                self._dgrad(la, dav, C[l], N, Hl, Wl, dp, C[l - 1])

    # -- training ----------------------------------------------------------
    def mse(self, out, target, mask=None):
        """(loss float64 tensor, dout bf16) of mean((out[..., 0] - target)^2 mask)."""
        N, H, W, _ = out.shape
        P = N * H * W
        t = target.to(self.device, torch.float32).contiguous()
        m = None if mask is None else mask.to(self.device, torch.float32).contiguous()
        count = float(P) if m is None else max(float(m.sum()), 1.0)
        dout = self._empty(N, H, W, IN_CH)
        loss = torch.empty((), dtype=torch.float64, device=self.device)
        part = torch.empty(_core.MISFIT_PARTIALS, dtype=torch.float64, device=self.device)
        _call("swp_nn_mse", _p(out), IN_CH, _p(t), _p(m), P, count, _p(dout), IN_CH, _p(loss), _p(part))
        return loss, dout

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_count += 1
        for lay in self.layers.values():
            _call("swp_nn_adam", _p(lay.w), _p(lay.gw), _p(lay.mw), _p(lay.vw), _p(lay.wbf), lay.w.numel(),
                  lr, b1, b2, eps, self.step_count)
            _call("swp_nn_adam", _p(lay.b), _p(lay.gb), _p(lay.mb), _p(lay.vb), None, lay.b.numel(),
                  lr, b1, b2, eps, self.step_count)

    def train_step(self, x, target, mask=None, lr=1e-3):
        out = self.forward(x)
        loss, dout = self.mse(out, target, mask)
        self.backward(dout)
        self.adam_step(lr)
        return loss


def conv_flops(net: UNet, N: int, H: int, W: int, training: bool = True) -> float:
    """Multiply-add flops (2 per MAC) of the convolutions of one forward (and,
    with training, the data and weight gradients; the first layer has no data
    gradient)."""
    total = 0.0
    for name, lay in net.layers.items():
        lvl = int(name[1]) if name[0] in "ed" else 0
        P = N * (H >> lvl) * (W >> lvl)
        f = 2.0 * P * lay.cin * lay.cout * lay.ks * lay.ks
        total += f
        if training:
            total += f * (1 if name == "e0a" else 2)
    return total


def _bench():
    """Training-step time of the U-Net at the 2-D workload's grid."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=2)
    ap.add_argument("--H", type=int, default=384)
    ap.add_argument("--W", type=int, default=1152)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    net = UNet(dev, levels=args.levels, base=args.base, nsrc=3)
    g = torch.Generator(device=dev).manual_seed(0)
    planes = [torch.randn((args.N, args.H, args.W), device=dev, generator=g) for _ in range(3)]
    x, _ = net.input(planes, [1.0, 1.0, 1.0])
    target = torch.randn((args.N, args.H, args.W), device=dev, generator=g)
    for _ in range(2):
        net.train_step(x, target)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.train_step(x, target)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    fl = conv_flops(net, args.N, args.H, args.W)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        net.forward(x)
    torch.cuda.synchronize()
    ms_f = (time.perf_counter() - t0) / args.steps * 1e3
    ff = conv_flops(net, args.N, args.H, args.W, training=False)
    print(json.dumps({"N": args.N, "H": args.H, "W": args.W, "levels": args.levels, "base": args.base,
                      "train_step_ms": round(ms, 3), "train_conv_tflops": round(fl / ms * 1e-9, 2),
                      "forward_ms": round(ms_f, 3), "forward_conv_tflops": round(ff / ms_f * 1e-9, 2),
                      "conv_gflop_per_train_step": round(fl * 1e-9, 2)}), flush=True)


if __name__ == "__main__":
    _bench()
