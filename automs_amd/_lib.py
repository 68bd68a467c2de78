"""ctypes binding of libams.so (include/ams.h).

The library is the only compute path: if it is missing or fails to load this
module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libams.so"

AMS_OK = 0
AMS_ERR_ARG = -1
AMS_ERR_HIP = -2
AMS_ERR_ALLOC = -3
AMS_ERR_OVERFLOW = -4
AMS_ADAM_STATE_BYTES = 32


class AmsArray(C.Structure):
    _fields_ = [
        ("n0", C.c_int32),
        ("n1", C.c_int32),
        ("pos", C.POINTER(C.c_float)),
        ("dir", C.c_float * 3),
        ("bptype", C.c_int32),
        ("radius", C.c_float),
    ]


class AmsTraceOpts(C.Structure):
    _fields_ = [
        ("wave_len", C.c_double),
        ("max_reflection", C.c_int32),
        ("gridsize", C.c_int32),
        ("fib_n", C.c_int64),
        ("per_element", C.c_int32),
        ("flags", C.c_int32),
        ("tasks_per_unit", C.c_int32),
        ("stream", C.c_void_p),
        ("keys_out", C.POINTER(C.c_uint64)),
        ("keys_cap", C.c_int64),
        ("seed_keys", C.POINTER(C.c_uint64)),
        ("n_seed", C.c_int64),
        ("discovery_mode", C.c_int32),
        ("launch_x", C.c_int32),
        ("launch_y", C.c_int32),
        ("samples_per_launch", C.c_int32),
        ("reserve_size", C.c_int32),
        ("task_info_path", C.c_char_p),
    ]


STAT_INT_FIELDS = [
    "rays", "records_raw", "records_unique", "paths", "tasks", "rounds", "chains", "successes", "overflow",
    "fwd_queries", "fwd_nodes", "fwd_tris", "fwd_spheres", "bwd_queries", "bwd_nodes", "bwd_tris",
    "n_discover_launches", "n_evaluate_launches",
]
STAT_DBL_FIELDS = ["t_discover_ms", "t_evaluate_ms", "t_total_ms"]
STAT_TAIL_FIELDS = ["bwd_walks", "beam_pairs", "beam_nodes", "beam_tris", "beam_walk_legs", "beam_cands"]


class AmsTraceStats(C.Structure):
    _fields_ = ([(n, C.c_int64) for n in STAT_INT_FIELDS] + [(n, C.c_double) for n in STAT_DBL_FIELDS] +
                [(n, C.c_int64) for n in STAT_TAIL_FIELDS])

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in STAT_INT_FIELDS + STAT_DBL_FIELDS + STAT_TAIL_FIELDS}


_lib = None


def lib() -> C.CDLL:
    """Load libams.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m automs_amd.build` "
            "(the HIP library is the only compute path)")
    L = C.CDLL(str(LIB_PATH))
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    L.ams_scene_create.argtypes = [fp, C.c_int32, ip, C.c_int32, fp, fp, C.POINTER(C.c_void_p)]
    L.ams_scene_create.restype = C.c_int
    L.ams_scene_destroy.argtypes = [C.c_void_p]
    L.ams_scene_destroy.restype = C.c_int
    L.ams_scene_plane_ids.argtypes = [C.c_void_p, ip]
    L.ams_scene_plane_ids.restype = C.c_int
    L.ams_trace_channel.argtypes = [C.c_void_p, C.POINTER(AmsArray), C.POINTER(AmsArray),
                                    C.POINTER(AmsTraceOpts), C.c_void_p, C.POINTER(AmsTraceStats)]
    L.ams_trace_channel.restype = C.c_int
    L.ams_last_error.argtypes = []
    L.ams_last_error.restype = C.c_char_p
    L.ams_version.restype = C.c_int
    L.ams_mem_outstanding.restype = C.c_int64
    _bind_contract(L)
    _lib = L
    return L


def _bind_contract(L):
    """Contraction, objective and Adam entry points (pointers as void*)."""
    vp, i32, i64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_float
    sigs = {
        "ams_frag_bytes": ([i32, i32], i64),
        "ams_frag_pack": ([vp, i32, i32, i64, i64, vp, vp], C.c_int),
        "ams_cgemm_workspace_bytes": ([i32, i32, i32], i64),
        "ams_cgemm_frag": ([vp, vp, i32, i32, i32, i32, vp, i64, i32, vp, vp], C.c_int),
        "ams_cgemm_f32": ([vp, i64, vp, i64, i32, i32, i32, i32, vp, i64, vp], C.c_int),
        "ams_cgemm_noalloc_max_bytes": ([], i64),
        "ams_valu_probe": ([vp, i32, i32, i32, vp], C.c_int),
        "ams_cgemm_frag_noalloc": ([vp, vp, i32, i32, i32, i32, vp, i64, i32, vp], C.c_int),
        "ams_objective_workspace_bytes": ([i32, i32], i64),
        "ams_objective_fwd": ([vp, i32, i32, i32, i32, f32, f32, vp, vp, vp, vp], C.c_int),
        "ams_objective_fwd_cov": ([vp, i32, i32, i32, i32, f32, f32, vp, vp, vp, vp, vp], C.c_int),
        "ams_objective_bwd": ([vp, i32, i32, i32, i32, f32, f32, vp, vp, vp, vp, vp], C.c_int),
        "ams_pm_operand": ([vp, vp, vp, vp, vp, i32, i32, i32, i32, f32, vp, vp, vp, vp], C.c_int),
        "ams_pm_operand_bwd": ([vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp], C.c_int),
        "ams_pm_operand_frag": ([vp, vp, vp, vp, vp, i32, i32, i32, i32, f32, vp, vp, vp, vp], C.c_int),
        "ams_objective_bwd_frag": ([vp, i32, i32, i32, i32, f32, f32, vp, vp, vp, vp, vp], C.c_int),
        "ams_chain_first": ([vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp], C.c_int),
        "ams_chain_step": ([vp, vp, i32, i32, i32, vp, i64, vp, vp], C.c_int),
        "ams_chain_bwd_scale": ([vp, i64, vp, vp, i32, i32, i32, vp, vp], C.c_int),
        "ams_chain_bwd_ms": ([vp, i64, vp, vp, i32, vp, i32, i32, i32, vp, i32, vp, vp, vp], C.c_int),
        "ams_chain_bwd_ap": ([vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp], C.c_int),
        "ams_adam_step": ([vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp], C.c_int),
        "ams_adam_step_dev": ([vp, vp, vp, vp, i64, f32, f32, f32, f32, vp, vp, vp, vp, i64, vp], C.c_int),
        "ams_host_bvh_build": ([vp, vp, i32, i32, vp, i32, vp, vp], C.c_int),
        "ams_host_bvh4_build": ([vp, vp, i32, i32, vp, i32, vp, vp], C.c_int),
        "ams_conv3x3_fwd": ([vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp], C.c_int),
        "ams_conv3x3_bwd_weight": ([vp, vp, vp, vp, i32, i32, i32, i32, i32, vp], C.c_int),
        "ams_relu_bwd": ([vp, vp, i64, vp], C.c_int),
        "ams_conv3x3_wfrag_bytes": ([i32, i32], i64),
        "ams_conv3x3_pack_weights": ([vp, i32, i32, vp, vp], C.c_int),
        "ams_conv3x3_fwd_mfma": ([vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp], C.c_int),
        "ams_conv3x3_dw_workspace_bytes": ([i32, i32, i32, i32, i32], i64),
        "ams_conv3x3_bwd_weight_mfma": ([vp, vp, vp, vp, i32, i32, i32, i32, i32, vp, vp], C.c_int),
        "ams_mse_partials": ([i64], i64),
        "ams_mse": ([vp, vp, vp, i64, vp, vp, vp], C.c_int),
        "ams_to_bf16": ([vp, vp, i64, vp], C.c_int),
        "ams_host_plane_ids": ([vp, i32, vp], C.c_int),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def check(rc: int, what: str) -> None:
    if rc != AMS_OK:
        msg = lib().ams_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")


EXPORTED = [
    "ams_scene_create", "ams_scene_destroy", "ams_scene_plane_ids", "ams_trace_channel",
    "ams_last_error", "ams_version", "ams_mem_outstanding",
    "ams_frag_bytes", "ams_frag_pack", "ams_cgemm_workspace_bytes", "ams_cgemm_frag", "ams_cgemm_f32",
    "ams_objective_workspace_bytes", "ams_objective_fwd", "ams_objective_bwd", "ams_adam_step",
    "ams_adam_step_dev", "ams_objective_fwd_cov", "ams_pm_operand", "ams_pm_operand_bwd", "ams_pm_operand_frag",
    "ams_objective_bwd_frag",
    "ams_host_bvh_build", "ams_host_bvh4_build", "ams_host_plane_ids",
    "ams_conv3x3_fwd", "ams_conv3x3_bwd_weight", "ams_relu_bwd", "ams_mse_partials", "ams_mse", "ams_to_bf16",
    "ams_cgemm_noalloc_max_bytes", "ams_cgemm_frag_noalloc", "ams_valu_probe",
    "ams_chain_first", "ams_chain_step", "ams_chain_bwd_scale", "ams_chain_bwd_ms", "ams_chain_bwd_ap",
    "ams_conv3x3_wfrag_bytes", "ams_conv3x3_pack_weights", "ams_conv3x3_fwd_mfma", "ams_conv3x3_dw_workspace_bytes",
    "ams_conv3x3_bwd_weight_mfma",
]
