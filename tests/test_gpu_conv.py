"""BF16 convolution kernels (forward, backward-data, backward-weight and the layout helpers)
against float64 CPU references of the same operations on the same bf16 operands.

The products of two bf16 values are exact in float32, so the only difference to the float64
reference is float32 accumulation: the tolerances below are a few float32 ulps of the sum
of |terms|. No reference counterpart (the convolutions serve the learned warm start).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from conftest import announce

pytestmark = pytest.mark.gpu


def _nat():
    from galaxy_collision_amd import _native as nat

    return nat


def bf16_rand(shape, gen, scale=1.0):
    return (torch.randn(shape, generator=gen) * scale).to(torch.bfloat16)


def to_ncdhw(x_cl: torch.Tensor, r: int) -> torch.Tensor:
    """channels-last [r^3, C] -> [1, C, r, r, r] (float64)."""
    c = x_cl.shape[1]
    return x_cl.double().view(r, r, r, c).permute(3, 0, 1, 2).unsqueeze(0)


def to_cl(y: torch.Tensor) -> torch.Tensor:
    """[1, C, r, r, r] -> [r^3, C]."""
    c = y.shape[1]
    return y[0].permute(1, 2, 3, 0).reshape(-1, c)


def concat_input(xa, ca, xb, cb, r, up):
    a = to_ncdhw(xa, r // 2 if up else r)
    if up:
        a = a.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    if cb:
        return torch.cat([a, to_ncdhw(xb, r)], dim=1)
    return a


def pack(w32: torch.Tensor, cout: int, cin: int, transpose: int, dev) -> torch.Tensor:
    nat = _nat()
    n = ctypes.c_int64()
    nat.call("pm_conv3d_packed_size", cout, cin, transpose, ctypes.byref(n))
    out = torch.empty(n.value, dtype=torch.bfloat16, device=dev)
    nat.call("pm_conv3d_pack_weights", nat.ptr(w32), cout, cin, transpose,
             nat.ptr(out, torch.bfloat16), nat.stream())
    return out


def store(c):
    return (c + 7) // 8 * 8


def scratch_for(cin, cout, r, split, dev):
    """Split-K scratch (or None): small grids then reduce over several workgroups."""
    if not split:
        return None, 0
    nat = _nat()
    n = ctypes.c_int64()
    nat.call("pm_conv3d_scratch_size", cin, cout, r, ctypes.byref(n))
    buf = torch.empty(max(n.value, 1), dtype=torch.float32, device=dev)
    return buf, n.value


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


FWD_CASES = [
    # r, ca, cb, up, cout, act
    (8, 8, 0, 0, 16, 1),
    (6, 16, 0, 0, 16, 0),     # 216 voxels: partial waves
    (16, 16, 0, 0, 32, 1),
    (8, 32, 0, 0, 64, 1),
    (4, 64, 0, 0, 128, 1),
    (8, 32, 16, 1, 16, 1),    # decoder: upsampled xa ++ xb
    (8, 128, 64, 1, 64, 1),
    (8, 16, 0, 0, 1, 0),      # output layer
    (4, 16, 0, 0, 192, 0),    # more than 128 output channels (two launches)
    (8, 8, 0, 0, 48, 1),      # 3 column blocks
]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("r,ca,cb,up,cout,act", FWD_CASES)
def test_conv_forward_matches_float64(device, r, ca, cb, up, cout, act, split):
    announce(f"conv fwd r={r} ca={ca} cb={cb} up={up} cout={cout} act={act} split={split}")
    nat = _nat()
    g = torch.Generator().manual_seed(100 + r * 7 + ca + cout)
    ra = r // 2 if up else r
    xa = bf16_rand((ra ** 3, ca), g)
    xb = bf16_rand((r ** 3, cb), g) if cb else None
    cin = ca + cb
    w32 = torch.randn(cout, cin, 3, 3, 3, generator=g) / np.sqrt(27 * cin)
    wb = w32.to(torch.bfloat16)
    bias = torch.randn(cout, generator=g) * 0.1
    slope = 0.1
    xa_d, w_d, b_d = xa.to(device), w32.to(device), bias.to(device)
    xb_d = xb.to(device) if cb else None
    packed = pack(w_d, cout, cin, 0, device)
    y = torch.empty(r ** 3 * store(cout), dtype=torch.bfloat16, device=device)
    y32 = torch.empty(r ** 3 * cout, dtype=torch.float32, device=device)
    scr, nscr = scratch_for(cin, cout, r, split, device)
    if split:
        assert nscr > 0, "these small grids split their reduction"
    nat.call("pm_conv3d_fwd_bf16", nat.ptr(xa_d, torch.bfloat16), ca,
             nat.ptr(xb_d, torch.bfloat16) if cb else None, cb, r, up,
             nat.ptr(packed, torch.bfloat16), nat.ptr(b_d), cout, act, slope,
             nat.ptr(y, torch.bfloat16), nat.ptr(y32), nat.ptr(scr), nscr, nat.stream())
    torch.cuda.synchronize()
    x = concat_input(xa, ca, xb, cb, r, up)
    ref = F.conv3d(x, wb.double(), bias.double(), padding=1)
    if act:
        ref = torch.where(ref > 0, ref, ref * slope)
    ref = to_cl(ref).numpy()
    # bound of |sum| terms for the accumulation tolerance
    bound = to_cl(F.conv3d(x.abs(), wb.double().abs(), None, padding=1)).numpy() + 1.0
    got = y32.cpu().numpy().reshape(r ** 3, cout)
    err = np.abs(got - ref)
    assert np.all(err <= 1e-5 * bound), f"max err {err.max():.3e}"
    assert rel_l2(got, ref) < 1e-5
    # bf16 output is the rounded float32 output, pad channels zero
    yb = y.cpu().view(r ** 3, store(cout))
    assert torch.equal(yb[:, :cout], torch.from_numpy(got).to(torch.bfloat16))
    assert torch.count_nonzero(yb[:, cout:].float()) == 0


BWD_CASES = [
    # r, ca, cb, up, cout
    (8, 8, 0, 0, 16),
    (6, 16, 0, 0, 16),
    (8, 16, 0, 0, 32),
    (4, 64, 0, 0, 128),
    (8, 32, 16, 1, 16),
    (4, 128, 64, 1, 64),
    (8, 16, 0, 0, 1),
]


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("r,ca,cb,up,cout", BWD_CASES)
def test_conv_backward_matches_float64_autograd(device, r, ca, cb, up, cout, split):
    announce(f"conv bwd r={r} ca={ca} cb={cb} up={up} cout={cout} split={split}")
    nat = _nat()
    g = torch.Generator().manual_seed(7 + r + ca * 3 + cout)
    ra = r // 2 if up else r
    cin = ca + cb
    xa = bf16_rand((ra ** 3, ca), g)
    xb = bf16_rand((r ** 3, cb), g) if cb else None
    w32 = torch.randn(cout, cin, 3, 3, 3, generator=g) / np.sqrt(27 * cin)
    wb = w32.to(torch.bfloat16)
    dz = torch.zeros(r ** 3, store(cout), dtype=torch.bfloat16)
    dz[:, :cout] = bf16_rand((r ** 3, cout), g)

    # float64 reference: autograd of sum(conv(x, w) * dz) on the same bf16 operands
    x = concat_input(xa, ca, xb, cb, r, up).clone().requires_grad_(True)
    w = wb.double().clone().requires_grad_(True)
    b = torch.zeros(cout, dtype=torch.float64, requires_grad=True)
    y = F.conv3d(x, w, b, padding=1)
    (y * to_ncdhw(dz[:, :cout].contiguous(), r)).sum().backward()
    ref_dx = to_cl(x.grad).numpy()
    ref_dw = w.grad.numpy().reshape(cout, cin, 27)
    ref_db = b.grad.numpy()

    xa_d = xa.to(device)
    xb_d = xb.to(device) if cb else None
    dz_d = dz.reshape(-1).to(device)
    w_d = w32.to(device)
    packed_t = pack(w_d, cout, cin, 1, device)
    dx = torch.empty(r ** 3 * cin, dtype=torch.float32, device=device)
    scr, nscr = scratch_for(cin, cout, r, split, device)
    nat.call("pm_conv3d_bwd_data_bf16", nat.ptr(dz_d, torch.bfloat16), cout, r,
             nat.ptr(packed_t, torch.bfloat16), cin, nat.ptr(dx), nat.ptr(scr), nscr,
             nat.stream())
    nfl = ctypes.c_int64()
    nat.call("pm_conv3d_wgrad_partial_size", cin, cout, r, ctypes.byref(nfl))
    part = torch.empty(nfl.value, dtype=torch.float32, device=device)
    dw = torch.full((cout * cin * 27,), 5.0, dtype=torch.float32, device=device)
    db = torch.full((cout,), 5.0, dtype=torch.float32, device=device)
    nat.call("pm_conv3d_bwd_weight_bf16", nat.ptr(xa_d, torch.bfloat16), ca,
             nat.ptr(xb_d, torch.bfloat16) if cb else None, cb, r, up,
             nat.ptr(dz_d, torch.bfloat16), cout, nat.ptr(dw), nat.ptr(db), 0, nat.ptr(part),
             nfl.value, nat.stream())
    torch.cuda.synchronize()
    got_dx = dx.cpu().numpy().reshape(r ** 3, cin)
    got_dw = dw.cpu().numpy().reshape(cout, cin, 27)
    assert rel_l2(got_dx, ref_dx) < 1e-5
    assert rel_l2(got_dw, ref_dw) < 1e-5
    np.testing.assert_allclose(db.cpu().numpy(), ref_db, rtol=1e-5, atol=1e-4)
    # accumulate adds
    nat.call("pm_conv3d_bwd_weight_bf16", nat.ptr(xa_d, torch.bfloat16), ca,
             nat.ptr(xb_d, torch.bfloat16) if cb else None, cb, r, up,
             nat.ptr(dz_d, torch.bfloat16), cout, nat.ptr(dw), nat.ptr(db), 1, nat.ptr(part),
             nfl.value, nat.stream())
    torch.cuda.synchronize()
    assert rel_l2(dw.cpu().numpy().reshape(cout, cin, 27), 2 * ref_dw) < 1e-5


def test_pool_act_upsum_helpers(device):
    announce("pool2 / act_bwd / upsum")
    nat = _nat()
    g = torch.Generator().manual_seed(3)
    r, c = 8, 16
    # device copies are bound to names: a temporary's memory could be reused by the next
    # allocation before the kernel that reads it runs
    x = bf16_rand((r ** 3, c), g)
    x_d = x.to(device)
    y = torch.empty((r // 2) ** 3 * c, dtype=torch.bfloat16, device=device)
    nat.call("pm_pool2_bf16", nat.ptr(x_d, torch.bfloat16), r, c, nat.ptr(y, torch.bfloat16),
             nat.stream())
    # float32 sum of the 8 children in order (dz, dy, dx) = bits (2, 1, 0), times 1/8
    xv = x.float().view(r // 2, 2, r // 2, 2, r // 2, 2, c)
    acc = torch.zeros(r // 2, r // 2, r // 2, c)
    for child in range(8):
        acc = acc + xv[:, child >> 2, :, (child >> 1) & 1, :, child & 1, :]
    assert torch.equal(y.cpu().view(-1, c), (acc * 0.125).reshape(-1, c).to(torch.bfloat16))

    g1 = torch.randn(r ** 3, c, generator=g)
    g2 = torch.randn((r // 2) ** 3, c, generator=g)
    act = bf16_rand((r ** 3, c), g)
    dz = torch.empty(r ** 3 * c, dtype=torch.bfloat16, device=device)
    g1_d, g2_d, act_d = g1.to(device), g2.to(device), act.to(device)
    nat.call("pm_act_bwd_bf16", nat.ptr(g1_d), c, nat.ptr(g2_d), nat.ptr(act_d, torch.bfloat16),
             0.1, r, c, c, nat.ptr(dz, torch.bfloat16), nat.stream())
    up = to_cl(to_ncdhw(g2, r // 2).repeat_interleave(2, 2).repeat_interleave(2, 3)
               .repeat_interleave(2, 4)).float()
    want = ((g1 + up * 0.125) * torch.where(act.float() > 0, 1.0, 0.1)).to(torch.bfloat16)
    assert torch.equal(dz.cpu().view(-1, c), want)

    dxf = torch.randn(r ** 3, 24, generator=g)
    ga = torch.empty((r // 2) ** 3 * 16, dtype=torch.float32, device=device)
    dxf_d = dxf.to(device)
    nat.call("pm_upsum_f32", nat.ptr(dxf_d), r, 24, 16, nat.ptr(ga), nat.stream())
    ref = F.avg_pool3d(to_ncdhw(dxf[:, :16].contiguous(), r), 2) * 8
    np.testing.assert_allclose(ga.cpu().numpy().reshape(-1, 16), to_cl(ref).numpy(), rtol=1e-6,
                               atol=1e-6)


def test_conv_argument_checks(device):
    nat = _nat()
    x = torch.zeros(8 ** 3 * 12, dtype=torch.bfloat16, device=device)
    with pytest.raises(nat.PMChainError, match="multiples of 8"):
        nat.call("pm_conv3d_fwd_bf16", nat.ptr(x, torch.bfloat16), 12, None, 0, 8, 0,
                 nat.ptr(x, torch.bfloat16), None, 16, 0, 0.0, nat.ptr(x, torch.bfloat16), None,
                 None, 0, nat.stream())
