"""Learned warm start for the multigrid Poisson solve (no reference counterpart).

The reference warm-starts each solve from the previous step's potential
(solver/multigrid.py:465-466). The solver here can instead start from the linear
extrapolation 2*phi(t-1) - phi(t-2) and add a learned correction: ``WarmStartNet`` is a
4-level 3D convolutional encoder-decoder (16 -> 32 -> 64 -> 128 channels, 3x3x3 kernels,
bf16 weights and activations, float32 accumulation) whose forward and backward passes are
the library's HIP kernels (pm_wsnet_*). It reads, at grid_size/pool resolution, the
pooled residual of the extrapolated guess, the previous potential increment and the
density, and predicts the pooled error of the extrapolated guess. The multigrid then
polishes to its own tolerance, so accuracy is unchanged; it also judges each prediction
and drops any that does not lower the residual.

Training pairs are produced by the solver itself (``pm_mg_solve_warm`` with pair outputs)
while a simulation runs; ``train`` fits the network with Adam on those pairs, computing
every gradient with the network's own backward kernels (torch only holds the buffers and
applies the optimizer update).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from galaxy_collision_amd import _native as nat

IN_CHANNELS = 3       # residual, previous increment, 4*pi*G*rho
STORED_CHANNELS = 8   # channels of the stored (bf16) network input
WEIGHTS_DIR = Path(__file__).with_name("weights")


class WarmStartNet:
    """The encoder-decoder on libpmchain's bf16 convolution kernels."""

    def __init__(self, res: int, width=(16, 32, 64, 128), slope: float = 0.1, seed: int = 0,
                 out_scale: float = 0.1):
        nat.require_device()
        self.res = int(res)
        self.width = tuple(int(w) for w in width)
        self.slope = float(slope)
        desc = nat.WsnetDesc(self.res, IN_CHANNELS, (ctypes.c_int32 * 4)(*self.width), self.slope)
        h = ctypes.c_void_p()
        nat.call("pm_wsnet_create", ctypes.byref(desc), ctypes.byref(h))
        self._h = h
        n = ctypes.c_int64()
        nat.call("pm_wsnet_num_params", self._h, ctypes.byref(n))
        self.num_params = int(n.value)
        self.layers = []
        for i in range(15):
            info = nat.WsnetLayer()
            nat.call("pm_wsnet_layer_info", self._h, i, ctypes.byref(info))
            self.layers.append({
                "name": info.name.decode(), "cin": info.cin_a + info.cin_b, "cin_a": info.cin_a,
                "cin_b": info.cin_b, "cout": info.cout, "res": info.res,
                "upsample_a": info.upsample_a, "act": info.act,
                "w_offset": int(info.w_offset), "b_offset": int(info.b_offset)})
        dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.params = torch.zeros(self.num_params, dtype=torch.float32, device=dev)
        self.grads = torch.zeros_like(self.params)
        self.out = torch.zeros(self.res ** 3, dtype=torch.float32, device=dev)
        self.init_params(seed, out_scale)

    # -- parameters -------------------------------------------------------------------
    def weight(self, i: int) -> torch.Tensor:
        L = self.layers[i]
        n = L["cout"] * L["cin"] * 27
        return self.params[L["w_offset"]: L["w_offset"] + n].view(L["cout"], L["cin"], 3, 3, 3)

    def bias(self, i: int) -> torch.Tensor:
        L = self.layers[i]
        return self.params[L["b_offset"]: L["b_offset"] + L["cout"]]

    def init_params(self, seed: int = 0, out_scale: float = 0.1) -> None:
        """He-normal weights for the leaky-ReLU layers, a small output layer, zero biases."""
        g = torch.Generator().manual_seed(int(seed))
        host = torch.zeros(self.num_params, dtype=torch.float32)
        for i, L in enumerate(self.layers):
            fan_in = L["cin"] * 27
            if i == 0:
                fan_in = IN_CHANNELS * 27  # the pad input channels are always zero
            std = math.sqrt(2.0 / ((1.0 + self.slope ** 2) * fan_in))
            if not L["act"]:
                std *= out_scale
            n = L["cout"] * L["cin"] * 27
            host[L["w_offset"]: L["w_offset"] + n] = torch.randn(n, generator=g) * std
        self.params.copy_(host)
        self.sync_params()

    def sync_params(self) -> None:
        nat.call("pm_wsnet_set_params", self._h, nat.ptr(self.params), nat.stream())

    # -- passes -----------------------------------------------------------------------
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x: bf16 [res^3 * 8] (channels-last) -> float32 [res^3] (this object's buffer)."""
        if x.numel() != self.res ** 3 * STORED_CHANNELS:
            raise ValueError(f"input has {x.numel()} elements, expected {self.res ** 3 * 8}")
        nat.call("pm_wsnet_forward", self._h, nat.ptr(x, torch.bfloat16), nat.ptr(self.out),
                 nat.stream())
        return self.out

    def backward(self, dout: torch.Tensor, accumulate: bool = False) -> torch.Tensor:
        nat.call("pm_wsnet_backward", self._h, nat.ptr(dout), nat.ptr(self.grads),
                 1 if accumulate else 0, nat.stream())
        return self.grads

    # -- persistence ------------------------------------------------------------------
    def save(self, path) -> None:
        np.savez(path, params=self.params.cpu().numpy(), res=self.res,
                 width=np.array(self.width), slope=self.slope)

    @classmethod
    def load(cls, path) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            net = cls(int(z["res"]), tuple(int(w) for w in z["width"]), float(z["slope"]))
            net.params.copy_(torch.from_numpy(z["params"]))
        net.sync_params()
        return net

    def close(self) -> None:
        if getattr(self, "_h", None):
            nat.call("pm_wsnet_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class PairSet:
    """(input, target) pairs produced by warm-started solves, resident on the device."""

    res: int
    inputs: list = field(default_factory=list)    # bf16 [res^3 * 8]
    targets: list = field(default_factory=list)   # float32 [res^3]

    def __len__(self) -> int:
        return len(self.inputs)

    def add(self, x: torch.Tensor, t: torch.Tensor) -> None:
        self.inputs.append(x)
        self.targets.append(t)

    def new_buffers(self):
        dev = torch.device("cuda", torch.cuda.current_device())
        return (torch.empty(self.res ** 3 * STORED_CHANNELS, dtype=torch.bfloat16, device=dev),
                torch.empty(self.res ** 3, dtype=torch.float32, device=dev))


class Adam:
    """Adam on a flat float32 parameter vector (plain torch arithmetic)."""

    def __init__(self, params: torch.Tensor, lr: float = 1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.p = params
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        self.m = torch.zeros_like(params)
        self.v = torch.zeros_like(params)
        self.t = 0

    def step(self, grad: torch.Tensor) -> None:
        self.t += 1
        self.m.mul_(self.b1).add_(grad, alpha=1 - self.b1)
        self.v.mul_(self.b2).addcmul_(grad, grad, value=1 - self.b2)
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.p.addcdiv_(mh, vh.sqrt_().add_(self.eps), value=-self.lr)


def relative_error(net: WarmStartNet, pairs: PairSet, idx) -> float:
    """Mean over pairs of ||net(x) - t|| / ||t||."""
    errs = []
    for i in idx:
        out = net.forward(pairs.inputs[i])
        t = pairs.targets[i]
        errs.append(float((out - t).norm() / t.norm().clamp_min(1e-30)))
    return float(np.mean(errs)) if errs else float("nan")


def train(net: WarmStartNet, pairs: PairSet, *, epochs: int = 10, batch: int = 4,
          lr: float = 1e-3, lr_final: float | None = None, clip: float = 1.0, seed: int = 0,
          val_idx=None, log=None) -> list[dict]:
    """Fit the network to the pairs (MSE on the normalized target) with Adam.

    Gradients come from pm_wsnet_backward, accumulated over each minibatch and clipped to a
    global norm of ``clip`` (the network has no normalization layers; unclipped Adam steps
    can diverge). The learning rate follows a cosine from ``lr`` to ``lr_final`` (default
    lr/10). Returns one record per epoch with the training loss and, if val_idx is given, the
    validation relative error.
    """
    rng = np.random.default_rng(seed)
    val = set(val_idx or [])
    train_idx = [i for i in range(len(pairs)) if i not in val]
    opt = Adam(net.params, lr=lr)
    lr_end = lr / 10.0 if lr_final is None else float(lr_final)
    total = max(1, epochs * math.ceil(len(train_idx) / batch))
    step = 0
    dout = torch.empty_like(net.out)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(train_idx)
        tot = 0.0
        for b0 in range(0, len(order), batch):
            chunk = order[b0: b0 + batch]
            for j, i in enumerate(chunk):
                out = net.forward(pairs.inputs[i])
                t = pairs.targets[i]
                diff = out - t
                tot += float((diff * diff).mean())
                torch.mul(diff, 2.0 / (diff.numel() * len(chunk)), out=dout)
                net.backward(dout, accumulate=j > 0)
            if clip > 0:
                gn = net.grads.norm()
                net.grads.mul_(torch.clamp(clip / (gn + 1e-12), max=1.0))
            opt.lr = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * step / total))
            step += 1
            opt.step(net.grads)
            net.sync_params()
        rec = {"epoch": ep, "loss": tot / max(1, len(order))}
        if val:
            rec["val_rel_err"] = relative_error(net, pairs, sorted(val))
        hist.append(rec)
        if log:
            log(rec)
    return hist


__all__ = ["WarmStartNet", "PairSet", "Adam", "train", "relative_error", "IN_CHANNELS",
           "STORED_CHANNELS", "WEIGHTS_DIR"]
