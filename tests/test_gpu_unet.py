"""The U-Net kernels on the GPU (through the C ABI) against a plain PyTorch
float32 reference on the CPU, fed the same bf16-rounded values.

* convolution, weight gradient and data gradient with float32 outputs: only
  the float32 summation order differs, bar 1e-5 relative L2;
* bf16 outputs: within one bf16 rounding of the float32 reference;
* pool / upsample / ReLU adjoints: bit-exact against the same float ops;
* the whole network, forward and backward, against a torch model that
  rounds to bf16 at the same points (bars stated at each check), and a few
  Adam steps lower the training loss."""
from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from sweep_amd import _core
from sweep_amd.unet import IN_CH, UNet, _call, _p

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BF = torch.bfloat16


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-300))


def bf(x):
    return x.to(BF).float()


def ref_conv(x, w, bias, ks):
    """x (N, H, W, cin) float, w (cout, ks*ks, cin) -> (N, H, W, cout) float32."""
    cout, _, cin = w.shape
    wt = w.reshape(cout, ks, ks, cin).permute(0, 3, 1, 2).contiguous()
    y = F.conv2d(x.permute(0, 3, 1, 2), wt, bias, padding=ks // 2)
    return y.permute(0, 2, 3, 1).contiguous()


CASES = [(1, 8, 20, 8, 32, 3), (2, 16, 136, 32, 64, 3), (1, 12, 140, 96, 32, 3), (1, 6, 64, 64, 200, 3),
         (2, 10, 130, 32, 8, 1), (1, 8, 256, 384, 128, 3)]


@pytest.mark.parametrize("N,H,W,cin,cout,ks", CASES)
def test_conv_forward_and_gradients(N, H, W, cin, cout, ks):
    print(f"[unet] conv N={N} H={H} W={W} cin={cin} cout={cout} ks={ks}", flush=True)
    g = torch.Generator().manual_seed(N * 1000 + H + cin)
    xs, os = cin + 8, cout + 16          # strided views
    xfull = bf(torch.randn((N, H, W, xs), generator=g))
    x = xfull[..., 8:8 + cin]
    w = bf(torch.randn((cout, ks * ks, cin), generator=g) / np.sqrt(cin * ks * ks))
    b = torch.randn(cout, generator=g)
    want = ref_conv(x, w, b, ks)
    xg = xfull.to(BF).to(DEV)
    wg = w.to(BF).to(DEV).contiguous()
    bg = b.to(DEV)
    out = torch.zeros((N, H, W, os), device=DEV)
    _call("swp_nn_conv_fwd", _p(xg) + 16, N, H, W, cin, xs, _p(wg), cout, ks, _p(bg), 0, _p(out), os, 1)
    got = out[..., :cout].cpu()
    assert rel(got, want) <= 1e-5
    assert bool(torch.all(out[..., cout:] == 0))
    # bf16 output with ReLU: one bf16 rounding of the float32 value
    outb = torch.zeros((N, H, W, cout), dtype=BF, device=DEV)
    _call("swp_nn_conv_fwd", _p(xg) + 16, N, H, W, cin, xs, _p(wg), cout, ks, _p(bg), 1, _p(outb), cout, 0)
    r = torch.relu(want)
    # one bf16 ulp, plus an absolute floor for values the float32 summation
    # order moves across zero (ReLU then keeps one side only)
    tol = r.abs() * 2.0 ** -7 + 1e-5 * float(want.abs().max())
    bad = int(((outb.float().cpu() - r).abs() > tol).sum())
    assert bad == 0, f"{bad} bf16 outputs off by more than one ulp"

    # gradients against torch autograd (float32)
    d = bf(torch.randn((N, H, W, cout), generator=g))
    xr = x.clone().requires_grad_(True)
    wr = w.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    (ref_conv(xr, wr, br, ks) * d).sum().backward()
    dg = d.to(BF).to(DEV)
    dw = torch.zeros_like(wg, dtype=torch.float32)
    db = torch.zeros(cout, device=DEV)
    n = int(_core.load().swp_nn_conv_wgrad_scratch(N, H, W, cin, cout, ks))
    scratch = torch.empty(n, device=DEV)
    _call("swp_nn_conv_wgrad", _p(xg) + 16, N, H, W, cin, xs, _p(dg), cout, cout, ks, _p(dw), _p(db), _p(scratch))
    assert rel(dw.cpu(), wr.grad) <= 1e-5
    assert rel(db.cpu(), br.grad) <= 1e-5
    wt = torch.empty((cin, ks * ks, cout), dtype=BF, device=DEV)
    _call("swp_nn_conv_flip", _p(wg), cout, cin, ks, _p(wt))
    dx = torch.zeros((N, H, W, cin), device=DEV)
    _call("swp_nn_conv_fwd", _p(dg), N, H, W, cout, cout, _p(wt), cin, ks, None, 0, _p(dx), cin, 1)
    assert rel(dx.cpu(), xr.grad) <= 1e-5


def test_pool_upsample_relu_adjoints_bit_exact():
    g = torch.Generator().manual_seed(4)
    N, H, W, C, S = 2, 8, 12, 16, 24
    x = bf(torch.randn((N, H, W, S), generator=g))
    xg = x.to(BF).to(DEV)
    p = torch.empty((N, H // 2, W // 2, C), dtype=BF, device=DEV)
    _call("swp_nn_pool2_fwd", _p(xg) + 16, S, N, H, W, C, _p(p), C)
    v = x[..., 8:8 + C]
    want = ((v[:, 0::2, 0::2] + v[:, 0::2, 1::2]) + (v[:, 1::2, 0::2] + v[:, 1::2, 1::2])) * 0.25
    ok = bool(torch.equal(p.float().cpu(), bf(want)))
    assert ok
    # pool adjoint accumulates 0.25 dout into din
    d = bf(torch.randn((N, H // 2, W // 2, C), generator=g))
    base = bf(torch.randn((N, H, W, C), generator=g))
    dg = base.to(BF).to(DEV)
    _call("swp_nn_pool2_bwd_acc", _p(d.to(BF).to(DEV)), C, N, H, W, C, _p(dg), C)
    up = d.repeat_interleave(2, 1).repeat_interleave(2, 2)
    ok = bool(torch.equal(dg.float().cpu(), bf(base + 0.25 * up)))
    assert ok
    # nearest upsample and its adjoint
    u = torch.empty((N, H, W, C), dtype=BF, device=DEV)
    _call("swp_nn_up2_fwd", _p(p), C, N, H // 2, W // 2, C, _p(u), C)
    ok = bool(torch.equal(u.float().cpu(), p.float().cpu().repeat_interleave(2, 1).repeat_interleave(2, 2)))
    assert ok
    du = bf(torch.randn((N, H, W, C), generator=g))
    back = torch.empty((N, H // 2, W // 2, C), dtype=BF, device=DEV)
    _call("swp_nn_up2_bwd", _p(du.to(BF).to(DEV)), C, N, H // 2, W // 2, C, _p(back), C)
    s = (du[:, 0::2, 0::2] + du[:, 0::2, 1::2]) + (du[:, 1::2, 0::2] + du[:, 1::2, 1::2])
    ok = bool(torch.equal(back.float().cpu(), bf(s)))
    assert ok
    # relu adjoint
    act = bf(torch.randn((N, H, W, C), generator=g))
    dd = bf(torch.randn((N, H, W, C), generator=g))
    ddg = dd.to(BF).to(DEV)
    _call("swp_nn_relu_bwd", _p(ddg), C, _p(act.to(BF).to(DEV)), C, N * H * W, C)
    ok = bool(torch.equal(ddg.float().cpu(), torch.where(act > 0, dd, torch.zeros_like(dd))))
    assert ok


class RefUNet:
    """torch float32 twin of sweep_amd.unet.UNet, rounding to bf16 where the
    GPU path stores bf16."""

    def __init__(self, net: UNet):
        self.net = net
        self.p = {n: (bf(l.w.detach().cpu()).requires_grad_(True), l.b.detach().cpu().clone().requires_grad_(True))
                  for n, l in net.layers.items()}

    def conv(self, name, x, relu=True, round_out=True):
        lay = self.net.layers[name]
        w, b = self.p[name]
        y = ref_conv(x, w, b, lay.ks)
        if relu:
            y = torch.relu(y)
        return _bf_ste(y) if round_out else y

    def forward(self, x):
        Ls, C = self.net.levels, self.net.C
        enc, inp = [], x
        for l in range(Ls):
            a = self.conv(f"e{l}a", inp)
            e = self.conv(f"e{l}b", a)
            enc.append(e)
            if l < Ls - 1:
                inp = _bf_ste(F.avg_pool2d(e.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))
        prev = enc[-1]
        for l in range(Ls - 2, -1, -1):
            u = prev.repeat_interleave(2, 1).repeat_interleave(2, 2)
            cat = torch.cat([u, enc[l]], dim=-1)
            prev = self.conv(f"d{l}b", self.conv(f"d{l}a", cat))
        return self.conv("out", prev, relu=False, round_out=False)


class _BfSte(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return bf(x)

    @staticmethod
    def backward(ctx, g):
        return bf(g)


def _bf_ste(x):
    return _BfSte.apply(x)


def test_unet_forward_backward_matches_torch():
    print("[unet] whole network 2 x 32 x 48, 3 levels, base 16", flush=True)
    net = UNet(DEV, levels=3, base=16, nsrc=3, seed=1)
    g = torch.Generator().manual_seed(9)
    N, H, W = 2, 32, 48
    planes = [torch.randn((N, H, W), generator=g) for _ in range(3)]
    scales = [1.0, 0.5, 2.0]
    x, _ = net.input([p.to(DEV) for p in planes], scales)
    xr = torch.zeros((N, H, W, IN_CH))
    for c in range(3):
        xr[..., c] = bf(planes[c] * scales[c])
    ok = bool(torch.equal(x.float().cpu(), xr))
    assert ok
    out = net.forward(x)
    ref = RefUNet(net)
    want = ref.forward(xr)
    e_fwd = rel(out[..., 0].cpu(), want[..., 0])
    print(f"[unet] forward rel {e_fwd:.2e}", flush=True)
    assert e_fwd <= 1e-4          # measured 1.4e-8: both round to bf16 at the same points
    ok = bool(torch.all(out[..., 1:] == 0))
    assert ok
    target = torch.randn((N, H, W), generator=g)
    loss, dout = net.mse(out, target.to(DEV))
    P = N * H * W
    r32 = (out[..., 0].cpu() - target).double()          # the residual in float32, as the kernel forms it
    lref = float((r32 ** 2).sum() / P)
    assert loss.item() == pytest.approx(lref, rel=1e-10)
    net.backward(dout)
    # reference gradient from the same bf16 dout
    want_out = want.clone()
    dref = dout.float().cpu()
    (want_out * dref).sum().backward()
    worst = 0.0
    for name, lay in net.layers.items():
        wr, br = ref.p[name]
        e = rel(lay.gw.cpu(), wr.grad)
        worst = max(worst, e)
        print(f"[unet] grad {name}: rel {e:.2e}", flush=True)
    assert worst <= 1e-2          # measured <= 2.4e-3: the backward's bf16 roundings of summed gradients


def test_unet_training_lowers_the_loss():
    print("[unet] training 20 Adam steps", flush=True)
    net = UNet(DEV, levels=4, base=16, nsrc=3, seed=2)
    g = torch.Generator().manual_seed(3)
    N, H, W = 2, 32, 64
    planes = [torch.randn((N, H, W), generator=g).to(DEV) for _ in range(3)]
    x, _ = net.input(planes, [1.0, 1.0, 1.0])
    target = (0.5 * planes[0] - 0.25 * planes[2]).contiguous()
    losses = [net.train_step(x, target, lr=3e-3).item() for _ in range(20)]
    print(f"[unet] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.5 * losses[0]


def test_save_load_resumes_training_exactly(tmp_path):
    """save() then load() into a fresh network of another seed restores the
    weights, biases and Adam moments: the next training step of both networks
    gives the same loss and the same forward output, bit for bit."""
    print("[unet] save / load round trip", flush=True)
    g = torch.Generator().manual_seed(5)
    N, H, W = 2, 32, 64
    planes = [torch.randn((N, H, W), generator=g).to(DEV) for _ in range(3)]
    a = UNet(DEV, levels=3, base=16, nsrc=3, seed=4)
    x, _ = a.input(planes, [1.0, 1.0, 1.0])
    target = (0.5 * planes[0] - 0.25 * planes[2]).contiguous()
    for _ in range(3):
        a.train_step(x, target, lr=3e-3)
    a.save(tmp_path / "net.pt")
    b = UNet(DEV, levels=3, base=16, nsrc=3, seed=9)
    assert not torch.equal(b.forward(x), a.forward(x))
    b.load(tmp_path / "net.pt")
    assert b.step_count == a.step_count
    la, lb = a.train_step(x, target, lr=3e-3).item(), b.train_step(x, target, lr=3e-3).item()
    assert la == lb
    assert torch.equal(a.forward(x), b.forward(x))
    with pytest.raises(ValueError):
        UNet(DEV, levels=3, base=8, nsrc=3, seed=0).load(tmp_path / "net.pt")
