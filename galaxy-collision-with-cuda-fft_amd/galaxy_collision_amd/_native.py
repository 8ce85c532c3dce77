"""ctypes binding of libpmchain.so (declared in include/pmchain.h).

The library is required: importing the package on a machine without the built library, or
calling a kernel without a HIP device, raises instead of falling back to any CPU path.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

LIB_PATH = Path(os.environ.get("PMCHAIN_LIB", Path(__file__).with_name("libpmchain.so")))

PM_OK = 0
PM_ERR_CAPACITY = 4
PM_MG_PLAIN = 0
PM_MG_MIXED = 1
PM_MG_OPT_GRAPHS = 0
PM_MG_OPT_SPLIT_FINE = 1
PM_MG_OPT_BOTTOM = 2
PM_MG_OPT_SMALL_SWEEP = 3
PM_DEPOSIT_SORTED = 0
PM_DEPOSIT_FIXED = 1
PM_DEPOSIT_BRICKED = 2
PM_STEP_BIN = 1
PM_STEP_RHO_KEPT = 2
PM_WARM_PREVIOUS = 0
PM_WARM_EXTRAPOLATE = 1
PM_WARM_LEARNED = 2

_vp = ctypes.c_void_p
_f = ctypes.c_float
_d = ctypes.c_double
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class PMChainError(RuntimeError):
    """An error status returned by libpmchain."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class MGDesc(ctypes.Structure):
    _fields_ = [
        ("grid_size", _i32),
        ("dx", _f),
        ("grav", _d),
        ("n_cycles", _i32),
        ("pre_sweeps", _i32),
        ("post_sweeps", _i32),
        ("coarse_sweeps", _i32),
        ("min_cycles", _i32),
        ("tol", _d),
        ("mode", _i32),
    ]


class MGStats(ctypes.Structure):
    _fields_ = [
        ("cycles", _i32),
        ("levels", _i32),
        ("resid_norm", _d),
        ("rhs_norm", _d),
        ("err_est", _d),
    ]


class MGWarm(ctypes.Structure):
    _fields_ = [
        ("mode", _i32),
        ("pool", _i32),
        ("net", _vp),
        ("fallback_ratio", _d),
        ("pair_input", _vp),
        ("pair_target", _vp),
        ("extrapolated", _i32),
        ("used_net", _i32),
        ("resid_before", _d),
        ("resid_after", _d),
    ]


class WsnetDesc(ctypes.Structure):
    _fields_ = [
        ("res", _i32),
        ("in_channels", _i32),
        ("width", _i32 * 4),
        ("slope", _f),
    ]


class WsnetLayer(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 8),
        ("cin_a", _i32),
        ("cin_b", _i32),
        ("cout", _i32),
        ("res", _i32),
        ("upsample_a", _i32),
        ("act", _i32),
        ("w_offset", _i64),
        ("b_offset", _i64),
    ]


class ParticleArrays(ctypes.Structure):
    """pm_particle_arrays: device pointers of one particle set's arrays."""

    _fields_ = [(k, ctypes.c_void_p) for k in
                ("pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "mass", "gid", "pid")]


class KernelTime(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48),
        ("launches", _i64),
        ("total_ms", _d),
        ("bytes", _d),
    ]


def timing_enable(on: bool | int) -> None:
    """False/0 off, True/1 every instrumented launch, 2 dispatch-recorded launches only."""
    call("pm_timing_enable", int(on))


def timing_reset() -> None:
    call("pm_timing_reset")


def timing_report() -> dict:
    """{kernel name: {"launches", "total_ms", "bytes"}} accumulated since the last reset."""
    count = _i32()
    call("pm_timing_report", 0, None, ctypes.byref(count))
    buf = (KernelTime * max(count.value, 1))()
    call("pm_timing_report", count.value, buf, ctypes.byref(count))
    return {k.name.decode(): {"launches": int(k.launches), "total_ms": float(k.total_ms),
                              "bytes": float(k.bytes)} for k in buf[: count.value]}


# name -> argtypes; every function returns pm_status (int)
_SIGNATURES = {
    "pm_deposit_plan_create": [_i32, _i64, ctypes.POINTER(_vp)],
    "pm_deposit_plan_destroy": [_vp],
    "pm_deposit_cic": [_vp, _vp, _vp, _vp, _vp, _i64, _f, _vp, _vp, _vp, _vp],
    "pm_deposit_cic_keyed": [_vp, _vp, _vp, _vp, _vp, _i64, _f, _vp, _vp, _vp, _vp],
    "pm_deposit_cic_step": [_vp, _vp, _vp, _vp, _vp, _i64, _f, _vp, _i32, _vp],
    "pm_deposit_plan_mark_ordered": [_vp],
    "pm_deposit_cic_noalloc": [_i32, _vp, _vp, _vp, _vp, _i64, _f, _vp, _vp, _vp, _vp],
    "pm_deposit_plan_set_timing": [_vp, _i32],
    "pm_deposit_plan_set_mode": [_vp, _i32],
    "pm_deposit_last_timing": [_vp, ctypes.POINTER(_f)],
    "pm_grad": [_vp, _vp, _vp, _vp, _i32, _f, _vp],
    "pm_gather": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _f, _vp],
    "pm_kick": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _f, _vp],
    "pm_drift": [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _f, _vp],
    "pm_kick_drift": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _f, _f, _vp],
    "pm_gather_kick": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32,
                       _f, _f, _vp],
    "pm_grad4": [_vp, _vp, _i32, _f, _vp],
    "pm_grad4_bricks": [_vp, _vp, _i32, _f, _vp, _vp],
    "pm_deposit_plan_node_bricks": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i32)],
    "pm_deposit_plan_order": [_vp, ctypes.POINTER(_vp)],
    "pm_deposit_plan_brick_list": [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                   ctypes.POINTER(_i64)],
    "pm_grad4_listed": [_vp, _vp, _i32, _f, _vp, _vp, _i64, _vp],
    "pm_gather_kick_drift4_reorder": [_vp, ctypes.POINTER(ParticleArrays),
                                      ctypes.POINTER(ParticleArrays), _vp, _i64, _i32, _f, _f,
                                      _f, _f, _vp, _vp],
    "pm_gather_kick4": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _f, _f,
                        _vp],
    "pm_gather_kick_drift4": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _f, _f, _f, _f,
                              _vp, _vp],
    "pm_direct_accel": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _f, _f, _vp],
    "pm_mg_create": [ctypes.POINTER(MGDesc), ctypes.POINTER(_vp)],
    "pm_mg_destroy": [_vp],
    "pm_mg_num_levels": [_vp, ctypes.POINTER(_i32)],
    "pm_mg_solve": [_vp, _vp, _vp, _i32, ctypes.POINTER(MGStats), _vp],
    "pm_mg_solve_warm": [_vp, _vp, _vp, ctypes.POINTER(MGWarm), ctypes.POINTER(MGStats), _vp],
    "pm_mg_warm_reset": [_vp],
    "pm_mg_vcycle": [_vp, _vp],
    "pm_mg_set_option": [_vp, _i32, _i32],
    "pm_mg_residual_norm": [_vp, ctypes.POINTER(_d), _vp],
    "pm_mg_moments": [_vp, _vp, ctypes.POINTER(_d), _vp],
    "pm_rb_pass": [_vp, _vp, _i32, _f, _i32, _vp],
    "pm_rb_sweep": [_vp, _vp, _vp, _i32, _f, _vp],
    "pm_residual": [_vp, _vp, _vp, _i32, _f, _vp],
    "pm_restrict": [_vp, _vp, _i32, _i32, _vp],
    "pm_prolong_add": [_vp, _vp, _i32, _i32, _vp],
    "pm_set_faces": [_vp, _i32, _f, ctypes.POINTER(_f), _vp],
    "pm_reduce_particles_f64": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, ctypes.POINTER(_d), _vp],
    "pm_reduce_grid_pe_f64": [_vp, _vp, _i64, ctypes.POINTER(_d), _vp],
    "pm_radial_histogram_f64": [_vp, _vp, _vp, _vp, _i64, _f, _f, _f, _f, _i32, _vp, _vp],
    "pm_copy_f32": [_vp, _vp, _i64, _vp],
    "pm_project_scratch_bytes": [_i32, ctypes.POINTER(_i64)],
    "pm_project_mass": [_vp, _vp, _vp, _i64, _f, _f, _f, _f, _i32, _vp, _vp, _vp],
    "pm_conv3d_packed_size": [_i32, _i32, _i32, ctypes.POINTER(_i64)],
    "pm_conv3d_pack_weights": [_vp, _i32, _i32, _i32, _vp, _vp],
    "pm_conv3d_scratch_size": [_i32, _i32, _i32, ctypes.POINTER(_i64)],
    "pm_conv3d_fwd_bf16": [_vp, _i32, _vp, _i32, _i32, _i32, _vp, _vp, _i32, _i32, _f, _vp, _vp,
                           _vp, _i64, _vp],
    "pm_conv3d_bwd_data_bf16": [_vp, _i32, _i32, _vp, _i32, _vp, _vp, _i64, _vp],
    "pm_conv3d_wgrad_partial_size": [_i32, _i32, _i32, ctypes.POINTER(_i64)],
    "pm_conv3d_bwd_weight_bf16": [_vp, _i32, _vp, _i32, _i32, _i32, _vp, _i32, _vp, _vp, _i32,
                                  _vp, _i64, _vp],
    "pm_pool2_bf16": [_vp, _i32, _i32, _vp, _vp],
    "pm_act_bwd_bf16": [_vp, _i32, _vp, _vp, _f, _i32, _i32, _i32, _vp, _vp],
    "pm_upsum_f32": [_vp, _i32, _i32, _i32, _vp, _vp],
    "pm_f32_to_bf16": [_vp, _i64, _vp, _vp],
    "pm_wsnet_create": [ctypes.POINTER(WsnetDesc), ctypes.POINTER(_vp)],
    "pm_wsnet_destroy": [_vp],
    "pm_wsnet_num_params": [_vp, ctypes.POINTER(_i64)],
    "pm_wsnet_layer_info": [_vp, _i32, ctypes.POINTER(WsnetLayer)],
    "pm_wsnet_set_params": [_vp, _vp, _vp],
    "pm_wsnet_forward": [_vp, _vp, _vp, _vp],
    "pm_wsnet_backward": [_vp, _vp, _vp, _i32, _vp],
    "pm_timing_enable": [_i32],
    "pm_timing_reset": [],
    "pm_timing_report": [_i32, _vp, ctypes.POINTER(_i32)],
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libpmchain.so once; raise if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"libpmchain.so not found at {LIB_PATH}: build it with "
                "`python galaxy-collision-with-cuda-fft_amd/build_native.py`")
        L = ctypes.CDLL(str(LIB_PATH))
        L.pm_last_error.restype = ctypes.c_char_p
        L.pm_last_error.argtypes = []
        L.pm_abi_version.restype = _i32
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def exported_names() -> list[str]:
    return ["pm_last_error", "pm_abi_version", *_SIGNATURES]


def call(name: str, *args) -> None:
    """Call a pm_* entry point and raise PMChainError on a non-zero status."""
    L = lib()
    status = getattr(L, name)(*args)
    if status != PM_OK:
        msg = L.pm_last_error().decode(errors="replace")
        raise PMChainError(status, f"{name}: {msg}")


def require_device() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("libpmchain kernels need a HIP device (torch.cuda.is_available() is False)")


def ptr(t: torch.Tensor | None, dtype: torch.dtype = torch.float32) -> int | None:
    """Device address of a contiguous HIP tensor (None passes through)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        t = getattr(t, "tensor", None)
        if t is None:
            raise TypeError("expected a torch tensor or a DeviceField")
    if not t.is_cuda:
        raise ValueError("libpmchain works on device tensors; got a host tensor")
    if t.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("libpmchain needs contiguous tensors")
    return t.data_ptr()


def stream(device: torch.device | None = None) -> int:
    """The HIP stream torch is currently issuing work on (hipStream_t as an integer)."""
    return torch.cuda.current_stream(device).cuda_stream
