"""ctypes binding of libmcphylo_hip.so (the C ABI in include/mcphylo_gpu.h).

There is no CPU fallback: if the library is missing, or no HIP device is
present, every call raises.  `build()` compiles the library in-tree with
hipcc for gfx950.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MCPHYLO_HIP_LIB selects another build of the same library (e.g. a tuned variant)
LIB_PATH = os.environ.get("MCPHYLO_HIP_LIB") or os.path.join(_HERE, "libmcphylo_hip.so")
CSRC = os.path.join(_HERE, "csrc")

MCP_OK = 0
MCP_ERR_INVALID = 1
MCP_ERR_DIMENSION = 2
MCP_ERR_COMPLEX = 3
MCP_ERR_NOMEM = 4
MCP_ERR_HIP = 5
MCP_ERR_UNSUPPORTED = 6
MCP_ERR_STATE = 7
MCP_FLAG_COMPRESS_PATTERNS = 1
MCP_SCHEDULE_FUSED = 0
MCP_SCHEDULE_LEVELS = 1
MCP_SCHEDULE_HYBRID = 2

# Every symbol the header declares (checked by the CPU test suite).
EXPORTED = (
    "mcp_plan_create", "mcp_plan_destroy", "mcp_set_model", "mcp_set_topology", "mcp_logl",
    "mcp_logl_grad", "mcp_batch_create", "mcp_batch_destroy", "mcp_batch_set_model",
    "mcp_batch_set_topology", "mcp_batch_logl", "mcp_batch_logl_grad", "mcp_use_graph",
    "mcp_set_schedule", "mcp_hmc_set_prior", "mcp_hmc_set_prior_params", "mcp_hmc_leapfrog",
    "mcp_set_tiles", "mcp_get_tiles", "mcp_set_hybrid_width", "mcp_autotune_schedule",
    "mcp_get_schedule", "mcp_get_launch_bytes",
    "mcp_autotune",
    "mcp_set_timing", "mcp_get_timing", "mcp_get_launch_times", "mcp_get_bytes", "mcp_get_sizes",
    "mcp_mlp_create", "mcp_mlp_destroy", "mcp_mlp_layout", "mcp_mlp_get_params",
    "mcp_mlp_set_params", "mcp_mlp_set_scaling", "mcp_mlp_set_data", "mcp_mlp_train",
    "mcp_mlp_param_grad", "mcp_mlp_predict", "mcp_mlp_leapfrog", "mcp_mlp_use_graph",
    "mcp_mlp_train_graph", "mcp_set_fused_waves",
    "mcp_logl_grad_by_rate", "mcp_batch_logl_grad_by_rate",
    "mcp_logl_grad_model", "mcp_batch_logl_grad_model",
    "mcp_nuts_create", "mcp_nuts_destroy", "mcp_nuts_use_batch", "mcp_nuts_use_mlp",
    "mcp_nuts_use_callback", "mcp_nuts_run", "mcp_nuts_get", "mcp_nuts_set",
    "mcp_last_error",
)

LEAPFROG_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double))


class McpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[mcp status {code}] {msg}")
        self.code = code


class McpDimensionMismatch(McpError, ValueError):
    """Mirrors Julia's DimensionMismatch."""


class McpArgumentError(McpError, ValueError):
    """Mirrors Julia's ArgumentError."""


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_fp = ctypes.POINTER(ctypes.c_float)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32), ("n_states", ctypes.c_int32),
        ("n_patterns", ctypes.c_int32), ("n_rates", ctypes.c_int32),
        ("parent", _ip), ("tip_partials", _dp), ("pattern_weights", _dp),
        ("device", ctypes.c_int32), ("flags", ctypes.c_uint32),
    ]


class DatasetDesc(ctypes.Structure):
    _fields_ = [
        ("n_states", ctypes.c_int32), ("n_patterns", ctypes.c_int32),
        ("n_leaves", ctypes.c_int32), ("leaf_nums", _ip), ("tip_partials", _dp),
        ("pattern_weights", _dp), ("flags", ctypes.c_uint32),
        ("tip_codes", ctypes.POINTER(ctypes.c_uint8)), ("n_codes", ctypes.c_int32),
        ("code_table", _dp),
    ]


class ItemDesc(ctypes.Structure):
    _fields_ = [
        ("dataset", ctypes.c_int32), ("n_nodes", ctypes.c_int32),
        ("n_rates", ctypes.c_int32), ("parent", _ip),
    ]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("n_datasets", ctypes.c_int32), ("datasets", ctypes.POINTER(DatasetDesc)),
        ("n_items", ctypes.c_int32), ("items", ctypes.POINTER(ItemDesc)),
        ("device", ctypes.c_int32), ("waves_per_block", ctypes.c_int32),
    ]


_lib = None


def build(verbose: bool = False) -> str:
    """Compile libmcphylo_hip.so for gfx950 (always recompiles)."""
    subprocess.run(["make", "-s", "-B", "-C", CSRC], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return LIB_PATH


def load_library():
    """Load the shared library without touching the GPU."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
            "(there is no CPU fallback for the likelihood path)")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "mcp_plan_create": (ctypes.c_int, [pp, ctypes.POINTER(PlanDesc)]),
        "mcp_plan_destroy": (None, [vp]),
        "mcp_set_model": (ctypes.c_int, [vp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]),
        "mcp_set_topology": (ctypes.c_int, [vp, _ip]),
        "mcp_logl": (ctypes.c_int, [vp, _dp, _dp]),
        "mcp_logl_grad": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "mcp_batch_create": (ctypes.c_int, [pp, ctypes.POINTER(BatchDesc)]),
        "mcp_batch_destroy": (None, [vp]),
        "mcp_batch_set_model": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp, _dp, ctypes.c_double,
                                               _dp, _dp]),
        "mcp_batch_set_topology": (ctypes.c_int, [vp, ctypes.c_int32, _ip]),
        "mcp_batch_logl": (ctypes.c_int, [vp, _dp, _dp]),
        "mcp_batch_logl_grad": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "mcp_use_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_set_schedule": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_hmc_set_prior": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]),
        "mcp_hmc_set_prior_params": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, _dp,
                                                     ctypes.c_int32]),
        "mcp_hmc_leapfrog": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "mcp_set_tiles": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
        "mcp_set_hybrid_width": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_autotune_schedule": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_int32, _ip]),
        "mcp_get_schedule": (ctypes.c_int, [vp, _ip, _ip]),
        "mcp_get_launch_bytes": (ctypes.c_int, [vp, _dp, ctypes.c_int32]),
        "mcp_get_tiles": (ctypes.c_int, [vp, ctypes.c_int32, _ip]),
        "mcp_autotune": (ctypes.c_int, [vp, ctypes.c_int32, _dp, ctypes.c_int32, _ip]),
        "mcp_set_timing": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_get_timing": (ctypes.c_int, [vp, _dp]),
        "mcp_get_launch_times": (ctypes.c_int, [vp, _dp, _ip, ctypes.c_int32]),
        "mcp_get_bytes": (ctypes.c_int, [vp, _dp]),
        "mcp_get_sizes": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "mcp_mlp_create": (ctypes.c_int, [pp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_uint64]),
        "mcp_mlp_destroy": (None, [vp]),
        "mcp_mlp_layout": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_int64)]),
        "mcp_mlp_get_params": (ctypes.c_int, [vp, _fp]),
        "mcp_mlp_set_params": (ctypes.c_int, [vp, _fp]),
        "mcp_mlp_set_scaling": (ctypes.c_int, [vp, _fp, _fp, ctypes.c_float, ctypes.c_float]),
        "mcp_mlp_set_data": (ctypes.c_int, [vp, _dp, _dp, _dp, ctypes.c_int64]),
        "mcp_mlp_train": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float, ctypes.c_uint64, _fp]),
        "mcp_mlp_param_grad": (ctypes.c_int, [vp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_float,
                                              _fp, _fp]),
        "mcp_mlp_predict": (ctypes.c_int, [vp, _dp, ctypes.c_int64, _dp, _dp]),
        "mcp_mlp_leapfrog": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double, _dp, _dp, _dp,
                                            _dp, _dp]),
        "mcp_mlp_use_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_mlp_train_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_set_fused_waves": (ctypes.c_int, [vp, ctypes.c_int32]),
        "mcp_logl_grad_by_rate": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "mcp_batch_logl_grad_by_rate": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "mcp_logl_grad_model": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "mcp_batch_logl_grad_model": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "mcp_nuts_create": (ctypes.c_int, [pp, ctypes.c_int32, _ip, _dp,
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.c_double,
                                           ctypes.c_int32, ctypes.c_int32]),
        "mcp_nuts_destroy": (None, [vp]),
        "mcp_nuts_use_batch": (ctypes.c_int, [vp, vp]),
        "mcp_nuts_use_mlp": (ctypes.c_int, [vp, vp, ctypes.c_double]),
        "mcp_nuts_use_callback": (ctypes.c_int, [vp, LEAPFROG_FN, vp]),
        "mcp_nuts_run": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                        ctypes.POINTER(ctypes.c_int64)]),
        "mcp_nuts_get": (ctypes.c_int, [vp, _dp, _dp, _dp, _ip]),
        "mcp_nuts_set": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "mcp_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    if status == MCP_OK:
        return
    msg = (load_library().mcp_last_error() or b"").decode(errors="replace")
    if status == MCP_ERR_DIMENSION:
        raise McpDimensionMismatch(status, msg)
    if status in (MCP_ERR_INVALID, MCP_ERR_UNSUPPORTED, MCP_ERR_COMPLEX):
        raise McpArgumentError(status, msg)
    raise McpError(status, msg)


def dptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def fptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_fp)


def device_count() -> int:
    """Number of HIP devices (0 when the runtime or a GPU is absent).  Uses
    the HIP runtime directly, not torch."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return 0
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return int(n.value)
