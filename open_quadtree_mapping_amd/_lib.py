"""ctypes binding of libqm.so (include/qm.h). No fallback: if the HIP library is
missing or fails to load, every call raises."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# QM_LIB names another build of the same library (the sanitizer build of the
# host-side checks); by default the in-tree libqm.so
LIB_PATH = os.environ.get("QM_LIB") or os.path.join(_HERE, "libqm.so")
_lib = None

P = C.c_void_p
I = C.c_int
F = C.c_float
D = C.c_double


STAGES = ("quadtree", "cost", "pyramid", "bp_sweep", "extract", "upsample_row", "upsample_col", "fuse_warp",
          "fuse_update")
TUNE_KNOBS = ("cost_waves", "bp_rows", "extract_rows", "overlap", "bp_pair", "bp_seg", "bp_pair_levels")  # QM_TUNE_* order
COST_REFERENCE, COST_FUSED = 0, 1  # QM_COST_*: qm_params.cost_arith
UPSAMPLE_THOMAS, UPSAMPLE_SCAN = 0, 1  # QM_UPSAMPLE_*: qm_params.upsample_solver


class QmError(RuntimeError):
    pass


class qm_params(C.Structure):
    _fields_ = [
        ("n_hyp", C.c_int),
        ("idepth_min", C.c_double),
        ("idepth_max", C.c_double),
        ("p1", C.c_float),
        ("p2", C.c_float),
        ("bp_iters", C.c_int * 4),
        ("lambda_", C.c_double),
        ("min_refs", C.c_int),
        ("max_refs", C.c_int),
        ("ref_cos", C.c_float),
        ("ref_baseline", C.c_float),
        ("cost_arith", C.c_int),
        ("upsample_solver", C.c_int),
    ]


class qm_frame_desc(C.Structure):
    _fields_ = [
        ("W", C.c_int),
        ("H", C.c_int),
        ("pitch", C.c_int),
        ("cam", C.c_float * 4),
        ("key", C.c_void_p),
        ("refs", C.c_void_p * 4),
        ("T_ref_key", C.c_float * 48),
        ("T_last_to_cur", C.c_float * 12),
        ("prefill", C.c_void_p),
        ("depth_out", C.c_void_p),
        ("dense_out", C.c_void_p),
        ("sparse_out", C.c_void_p),
        ("lambda_", C.c_double),
        ("prior", C.c_void_p),
        ("prior_depth", C.c_void_p),
    ]


class qm_upsample_problem(C.Structure):
    _fields_ = [
        ("W", C.c_int),
        ("H", C.c_int),
        ("row_seg", C.c_int),
        ("col_seg", C.c_int),
        ("lambda_", C.c_double),
        ("d", C.c_void_p),
        ("conf", C.c_void_p),
        ("w_row", C.c_void_p),
        ("w_col", C.c_void_p),
        ("dense", C.c_void_p),
        ("row_x", C.c_void_p),
        ("row_n", C.c_void_p),
        ("g", C.c_void_p),
        ("grad_d", C.c_void_p),
        ("grad_conf", C.c_void_p),
        ("grad_w_row", C.c_void_p),
        ("grad_w_col", C.c_void_p),
        ("grad_lambda", C.c_void_p),
        ("state", C.c_void_p),
    ]


ABI_VERSION = 13  # QM_ABI_VERSION of include/qm.h this binding is written against

# name -> (restype, argtypes); mirrors include/qm.h
SIGNATURES = {
    "qm_abi_version": (I, []),
    "qm_last_error": (C.c_char_p, []),
    "qm_default_params": (None, [P]),
    "qm_set_alloc_cap": (None, [C.c_size_t]),
    "qm_device_bytes_in_use": (C.c_size_t, []),
    "qm_release_scratch": (I, []),
    "qm_quadtree": (I, [P, I, I, I, P, P]),
    "qm_cost_volume": (I, [P, P, I, I, I, P, P, P, P, P, P, P]),
    "qm_bp_extract": (I, [P, I, I, P, P, P, P, P, P]),
    "qm_upsample": (I, [P, P, I, I, I, P, P, P]),
    "qm_upsample_general": (I, [P, P, P, P, I, I, D, I, I, P, P, P, P]),
    "qm_upsample_backward": (I, [P, P, P, P, I, I, D, I, I, P, P, P, P, P, P]),
    "qm_upsample_general_batch": (I, [P, I, P]),
    "qm_upsample_backward_batch": (I, [P, I, P]),
    "qm_fuse": (I, [P, P, I, I, P, P, P, P, P, P]),
    "qm_fuse_noalloc": (I, [P, P, I, I, P, P, P, P, P, P]),
    "qm_fuse_backward": (I, [P, P, P, I, I, P, P, P, P]),
    "qm_point_cloud": (I, [P, P, I, I, P, P, P, P, P]),
    "qm_upsample_local": (I, [P, P, I, I, I, P, P]),
    "qm_sd_gradient": (I, [P, I, I, I, P, P]),
    "qm_sd_init": (I, [P, I, I, P, P]),
    "qm_sd_step": (I, [P, P, P, P, P, I, I, I, P, P, P, I, P, P, P]),
    "qm_quadtree_prior": (I, [P, I, I, P, P]),
    "qm_prior_to_cost": (I, [P, I, I, P, P, P, P, P]),
    "qm_batch_create": (I, [P, I, P, P]),
    "qm_batch_set_frame": (I, [P, I, P]),
    "qm_batch_run": (I, [P, I, P]),
    "qm_batch_run_timed": (I, [P, P, P, P]),
    "qm_batch_buffer": (I, [P, I, I, P]),
    "qm_batch_read_buffer": (I, [P, I, I, P, C.c_size_t]),
    "qm_batch_reset_seeds": (I, [P]),
    "qm_batch_backward": (I, [P, P, P, P]),
    "qm_batch_runs": (I, [P, P]),
    "qm_batch_set_tuning": (I, [P, I, I]),
    "qm_batch_get_tuning": (I, [P, I, P]),
    "qm_batch_groups": (I, [P, P, P, P, I]),
    "qm_batch_set_group_tuning": (I, [P, I, I, I]),
    "qm_batch_get_group_tuning": (I, [P, I, I, P]),
    "qm_batch_autotune": (I, [P, I, P, P, P]),
    "qm_batch_destroy": (I, [P]),
    "qm_create": (I, [I, I, F, F, F, F, P, P]),
    "qm_destroy": (I, [P]),
    "qm_set_semi2dense_ratio": (I, [P, I]),
    "qm_set_semidense": (I, [P, I]),
    "qm_set_remap": (I, [P, P, P]),
    "qm_set_prior_alternates": (I, [P, I, I]),
    "qm_add_frame": (I, [P, P, I, I, I, P, P]),
    "qm_get_depth": (I, [P, P]),
    "qm_get_debug": (I, [P, P]),
    "qm_get_reference": (I, [P, P]),
    "qm_get_T_world_ref": (I, [P, P]),
    "qm_get_point_cloud": (I, [P, P, P]),
    "qm_num_references": (I, [P]),
}


def _init_torch_first():
    """torch carries its own HIP runtime. When both runtimes live in one
    process, torch's must initialise the device first (the other order leaves
    torch without a GPU), so it is started before libqm is loaded."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QmError(f"{LIB_PATH} is not built; run __graft_entry__.build() (the HIP path has no fallback)")
        _init_torch_first()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        got = L.qm_abi_version()
        if got != ABI_VERSION:
            raise QmError(f"{LIB_PATH} has ABI version {got}, the binding expects {ABI_VERSION}: rebuild it")
        _lib = L
    return _lib


def release_scratch():
    """Free the library's kept stage-level scratch (qm_release_scratch)."""
    check(lib().qm_release_scratch())


def device_empty(shape, dtype, device):
    """torch.empty on the device; when torch runs out of memory, the library's
    kept scratch is released and the allocation tried once more."""
    import torch
    try:
        return torch.empty(shape, dtype=dtype, device=device)
    except torch.cuda.OutOfMemoryError:
        release_scratch()
        return torch.empty(shape, dtype=dtype, device=device)


def check(rc: int):
    if rc != 0:
        msg = lib().qm_last_error()
        raise QmError(f"libqm error {rc}: {msg.decode() if msg else ''}")


def default_params(**kw) -> qm_params:
    p = qm_params()
    lib().qm_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "lambda":
            k = "lambda_"
        if k == "bp_iters":
            for i in range(4):
                p.bp_iters[i] = int(v[i])
            continue
        setattr(p, k, v)
    return p
