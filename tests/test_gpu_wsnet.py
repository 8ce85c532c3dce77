"""The warm-start network (pm_wsnet_*): forward pass and parameter gradients against a float64
CPU autograd restatement of the same network, and a short training run on the GPU kernels.

The restatement rounds every stored activation to bf16 exactly where the kernels do (layer
outputs, pooled tensors), with a straight-through rounding in the backward pass. The
kernels also round each layer's output gradient to bf16, which the float64 reference does
not. Where a float32 sum and its float64 counterpart fall on different sides of a bf16
rounding boundary the stored activation differs by one bf16 ulp (2^-8 relative); such flips
propagate through the 15 layers, so the output agrees to a few 1e-3 and the gradients to a
few 1e-2 (up to ~5e-2 for the first layer, the end of the backward chain); a wiring mistake
(wrong skip, pooling or upsampling gradient) shows as an O(1) error. The per-layer
primitives are pinned to float32 accuracy in test_gpu_conv.py.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from conftest import announce

pytestmark = pytest.mark.gpu


def bf(x: torch.Tensor) -> torch.Tensor:
    """Straight-through bf16 rounding of a float64 tensor."""
    return x + (x.float().to(torch.bfloat16).double() - x).detach()


def up2(x):
    return x.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)


def reference_forward(net, params64: torch.Tensor, x_in: torch.Tensor):
    """float64 restatement of the pm_wsnet forward pass; returns the output [R^3]."""
    L = net.layers
    slope = net.slope

    def conv(i, x):
        li = L[i]
        n = li["cout"] * li["cin"] * 27
        w = params64[li["w_offset"]: li["w_offset"] + n].view(li["cout"], li["cin"], 3, 3, 3)
        w = w + (w.float().to(torch.bfloat16).double() - w).detach()  # bf16 weights
        b = params64[li["b_offset"]: li["b_offset"] + li["cout"]]
        y = F.conv3d(x, w, b, padding=1)
        if li["act"]:
            y = torch.where(y > 0, y, y * slope)
        return y

    def pool(x):
        return bf(F.avg_pool3d(x, 2))

    ea0 = bf(conv(0, x_in))
    s0 = bf(conv(1, ea0))
    ea1 = bf(conv(2, pool(s0)))
    s1 = bf(conv(3, ea1))
    ea2 = bf(conv(4, pool(s1)))
    s2 = bf(conv(5, ea2))
    ea3 = bf(conv(6, pool(s2)))
    b3 = bf(conv(7, ea3))
    d2a = bf(conv(8, torch.cat([up2(b3), s2], 1)))
    d2b = bf(conv(9, d2a))
    d1a = bf(conv(10, torch.cat([up2(d2b), s1], 1)))
    d1b = bf(conv(11, d1a))
    d0a = bf(conv(12, torch.cat([up2(d1b), s0], 1)))
    d0b = bf(conv(13, d0a))
    return conv(14, d0b)[0, 0]


def make_input(res, gen, dev):
    x = torch.zeros(res ** 3, 8)
    x[:, :3] = torch.randn(res ** 3, 3, generator=gen)
    xb = x.to(torch.bfloat16)
    return xb.reshape(-1).to(dev), xb.double().view(res, res, res, 8).permute(3, 0, 1, 2)[None]


def test_network_forward_backward_match_float64(device):
    announce("wsnet fwd/bwd vs float64 autograd, res 16")
    from galaxy_collision_amd.warmstart import WarmStartNet

    res = 16
    net = WarmStartNet(res, seed=3, out_scale=1.0)
    g = torch.Generator().manual_seed(11)
    with torch.no_grad():  # non-zero biases so the bias path is exercised
        for i in range(15):
            net.bias(i).copy_(torch.randn(net.layers[i]["cout"], generator=g).to(device) * 0.05)
    net.sync_params()
    x_dev, x_ref = make_input(res, g, device)
    out = net.forward(x_dev).clone()
    dout = torch.randn(res ** 3, generator=g)
    net.backward(dout.to(device))
    torch.cuda.synchronize()

    p64 = net.params.detach().cpu().double().clone().requires_grad_(True)
    ref = reference_forward(net, p64, x_ref)
    (ref.reshape(-1) * dout.double()).sum().backward()
    got_out = out.cpu().double().numpy()
    ref_out = ref.detach().reshape(-1).numpy()
    rel = np.linalg.norm(got_out - ref_out) / np.linalg.norm(ref_out)
    assert rel < 1e-2, f"forward rel L2 {rel:.2e}"

    got_g = net.grads.cpu().double().numpy()
    ref_g = p64.grad.numpy()
    for i, L in enumerate(net.layers):
        n = L["cout"] * L["cin"] * 27
        sl = slice(L["w_offset"], L["w_offset"] + n)
        gw, rw = got_g[sl], ref_g[sl]
        if i == 0:  # the zero pad input channels carry no gradient
            keep = np.zeros((L["cout"], L["cin"], 27), bool)
            keep[:, :3] = True
            gw = gw.reshape(L["cout"], L["cin"], 27)[keep]
            rw = rw.reshape(L["cout"], L["cin"], 27)[keep]
        err = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        assert err < 0.1, f"layer {L['name']} weight grad rel L2 {err:.2e}"
        bsl = slice(L["b_offset"], L["b_offset"] + L["cout"])
        berr = np.linalg.norm(got_g[bsl] - ref_g[bsl]) / max(np.linalg.norm(ref_g[bsl]), 1e-30)
        assert berr < 0.1, f"layer {L['name']} bias grad rel L2 {berr:.2e}"
    cos = float(np.dot(got_g, ref_g) / (np.linalg.norm(got_g) * np.linalg.norm(ref_g)))
    assert cos > 0.999


def test_backward_accumulates(device):
    announce("wsnet backward accumulate")
    from galaxy_collision_amd.warmstart import WarmStartNet

    res = 8
    net = WarmStartNet(res, seed=5)
    g = torch.Generator().manual_seed(2)
    x_dev, _ = make_input(res, g, device)
    dout = torch.randn(res ** 3, generator=g).to(device)
    net.forward(x_dev)
    g1 = net.backward(dout).clone()
    net.backward(dout, accumulate=True)
    torch.cuda.synchronize()
    torch.testing.assert_close(net.grads, 2 * g1, rtol=1e-6, atol=1e-12)


def test_training_reduces_loss(device):
    """Adam on the network's own gradients learns a smooth linear map of its input."""
    announce("wsnet training smoke, res 16")
    from galaxy_collision_amd.warmstart import PairSet, WarmStartNet, train

    res = 16
    net = WarmStartNet(res, seed=1)
    g = torch.Generator().manual_seed(5)
    pairs = PairSet(res)
    for _ in range(8):
        x_dev, x_ref = make_input(res, g, device)
        # target: 3x3x3 box average of channel 0 minus half of channel 1
        t = F.avg_pool3d(F.pad(x_ref[:, :1], (1, 1, 1, 1, 1, 1)), 3, stride=1) - 0.5 * x_ref[:, 1:2]
        pairs.add(x_dev, t[0, 0].reshape(-1).float().to(device))
    hist = train(net, pairs, epochs=40, batch=4, lr=1e-3, seed=0)
    first, last = hist[0]["loss"], hist[-1]["loss"]
    assert np.isfinite(last) and last < 0.25 * first, (first, last)
