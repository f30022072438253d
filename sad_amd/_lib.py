"""ctypes binding of libsad_hip.so, the C ABI declared in include/sad_hip.h.

Every entry point returns an int status; `call` raises SadHipError with the
library's message when it is non-zero.  The library must be built (see
__graft_entry__.build or `make -C sad_amd/csrc`); there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SAD_HIP_LIBRARY names another build of the same library, e.g. a schedule variant built with
# `make -C sad_amd/csrc SCHEDULE=... OUT=...` for measurement; by default the in-tree build is loaded.
LIB_PATH = os.environ.get("SAD_HIP_LIBRARY") or os.path.join(_HERE, "libsad_hip.so")

P = C.c_void_p
U32 = C.c_uint32
I32 = C.c_int
F32 = C.c_float


class SadHipError(RuntimeError):
    pass


class FitOptions(C.Structure):
    """SadhFitOptions (include/sad_hip.h)."""
    _fields_ = [
        ("iterations", I32), ("target_bpp", F32), ("init_mode", I32), ("n_sites", I32), ("max_sites", I32),
        ("prune_percentile", F32), ("prune_start", I32), ("prune_freq", I32), ("prune_end", I32),
        ("prune_during_densify", I32), ("densify_enabled", I32), ("densify_start", I32), ("densify_freq", I32),
        ("densify_end", I32), ("densify_percentile", F32), ("densify_score_alpha", F32),
        ("cand_update_freq", I32), ("cand_update_passes", I32), ("cand_downscale", I32),
        ("cand_radius_scale", F32), ("cand_radius_probes", U32), ("cand_inject_count", U32),
        ("cand_hilbert_window", U32), ("cand_hilbert_probes", U32),
        ("lr_scale", F32), ("lr_pos", F32), ("lr_tau", F32), ("lr_radius", F32), ("lr_color", F32),
        ("lr_dir", F32), ("lr_aniso", F32), ("beta1", F32), ("beta2", F32), ("eps", F32),
        ("log_freq", I32), ("profile_gradients", I32), ("init_gradient_alpha", F32), ("init_log_tau", F32),
        ("init_radius", F32), ("ssim_weight", F32), ("ssim_metric", I32), ("packed_psnr", I32),
        ("graph_ticks", I32), ("autotune", I32),
    ]


class FitResult(C.Structure):
    """SadhFitResult (include/sad_hip.h)."""
    _fields_ = [
        ("final_psnr", F32), ("best_psnr", F32), ("actual_sites", I32), ("active_estimate", I32),
        ("active_count", I32), ("densify_percentile", F32), ("prune_percentile", F32), ("capacity", I32),
        ("n_logs", I32), ("n_densify", I32), ("n_prune", I32), ("n_cand_updates", I32),
        ("train_seconds", C.c_double), ("grad_kernel_ms", C.c_double), ("grad_launches", I32),
        ("grad_bytes", C.c_double), ("grad_ms_trace", C.POINTER(F32)), ("grad_ms_trace_cap", I32),
        ("final_render", C.POINTER(F32)), ("final_ssim", F32), ("best_ssim", F32), ("log_ssim", C.POINTER(F32)),
        ("packed_psnr", F32), ("packed_render", C.POINTER(F32)), ("cells_render", C.POINTER(F32)),
        ("tau_heatmap", C.POINTER(F32)),
    ]


def fit_result_dict(res: "FitResult") -> dict:
    """SadhFitResult fields as a dict, without its host-pointer fields."""
    skip = ("grad_ms_trace", "grad_ms_trace_cap", "final_render", "log_ssim", "packed_render", "cells_render",
            "tau_heatmap")
    return {k: getattr(res, k) for k, _ in FitResult._fields_ if k not in skip}


class BatchJob(C.Structure):
    """SadhBatchJob (include/sad_hip.h)."""
    _fields_ = [
        ("target", P), ("mask", P), ("width", I32), ("height", I32), ("sites_io", P), ("n_sites", I32),
        ("sites_cap", I32), ("log_psnr", P), ("log_cap", I32), ("result", C.POINTER(FitResult)),
    ]


class BatchResult(C.Structure):
    """SadhBatchResult (include/sad_hip.h)."""
    _fields_ = [
        ("train_seconds", C.c_double), ("grad_kernel_ms", C.c_double), ("grad_launches", I32),
        ("grad_bytes", C.c_double), ("total_pixels", C.c_double), ("graph_ticks", I32), ("graph_captures", I32),
        ("tuned_hash_slots", I32), ("tuned_tau_block", I32),
    ]


class ServerOptions(C.Structure):
    """SadhServerOptions (include/sad_hip.h)."""
    _fields_ = [("fit", FitOptions), ("memory_budget", C.c_double), ("max_fit_jobs", I32), ("render_passes", I32)]


class ServerEvent(C.Structure):
    """SadhServerEvent (include/sad_hip.h)."""
    _fields_ = [("id", C.c_longlong), ("kind", I32), ("status", I32), ("submit_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double)]


# name -> argument types (all return int)
SIGNATURES = {
    "sadh_render_voronoi": [P, P, P, P, F32, U32, I32, I32, I32, I32, P],
    "sadh_render_cells": [P, P, P, P, F32, U32, I32, I32, I32, I32, P],
    "sadh_render_tau_heatmap": [P, P, P, P, U32, F32, F32, F32, F32, I32, I32, I32, I32, P],
    "sadh_render_backward": [P, P, P, P, P, F32, U32, I32, I32, I32, I32, P],
    "sadh_init_candidates": [P, P, U32, U32, I32, I32, I32, P],
    "sadh_clear_candidates": [P, P, I32, I32, P],
    "sadh_jfa_seed": [P, P, U32, I32, I32, I32, P],
    "sadh_jfa_flood": [P, P, P, U32, U32, F32, I32, I32, I32, I32, I32, P],
    "sadh_pack_candidate_sites": [P, P, U32, P],
    "sadh_update_candidates": [P, P, P, P, P, U32, U32, F32, U32, F32, U32, U32, P, P, U32, U32,
                               I32, I32, I32, I32, I32, P],
    "sadh_hilbert_pairs": [P, P, U32, U32, I32, I32, I32, P],
    "sadh_hilbert_order": [P, P, P, U32, P],
    "sadh_compute_gradients_tiled": [P, P, P, P, P] + [P] * 10 + [P, F32, U32, P, U32, F32, I32, I32, I32, I32, P],
    "sadh_compute_gradients_tiled_noalloc": [P, P, P, P, P] + [P] * 10 + [P, F32, U32, P, U32, F32, I32, I32, I32,
                                                                         I32, P],
    "sadh_compute_gradients_tiled_aos": [P, P, P, P, P, P, F32, U32, P, U32, I32, I32, I32, I32, P],
    "sadh_tau_diffuse": [P, P, P, P, P, P, U32, F32, I32, I32, I32, P],
    "sadh_tau_diffuse_adam": [P, P, P, P] + [P] * 10 + [P, F32, I32, I32, I32] + [F32] * 9 + [U32, U32, I32, I32, P],
    "sadh_adam_update": [P, P] + [P] * 10 + [F32] * 9 + [U32, U32, I32, I32, P],
    "sadh_adam_update_clear": [P, P] + [P] * 10 + [F32] * 9 + [U32, U32, I32, I32, P],
    "sadh_clear_gradients": [P] * 10 + [U32, P],
    "sadh_clear_buffer": [P, U32, P],
    "sadh_compute_site_stats_simple": [P, P, P, P, P, F32, U32] + [P] * 8 + [I32, I32, I32, I32, P],
    "sadh_compute_densify_score_pairs": [P, P, P, P, U32, F32, F32, U32, P],
    "sadh_compute_prune_score_pairs": [P, P, P, U32, F32, U32, P],
    "sadh_radix_sort_pairs": [P, P, P, P, U32, U32, P],
    "sadh_write_split_indices": [P, P, U32, P],
    "sadh_split_sites": [P, P, P, U32] + [P] * 7 + [U32, P, I32, I32, P],
    "sadh_prune_sites_by_index": [P, P, U32, P],
    "sadh_compute_psnr": [P, P, P, P, I32, I32, P],
    "sadh_count_active_sites": [P, P, U32, P],
    "sadh_compute_ssim": [P, P, P, P, I32, I32, P],
    "sadh_pack_inference_sites": [P, U32, I32, I32, P, P, P],
    "sadh_unpack_inference_sites": [P, P, U32, I32, I32, P, P],
    "sadh_render_packed_sites": [P, P, U32, I32, I32, I32, I32, I32, I32, F32, U32, U32, U32, U32, P, C.POINTER(F32),
                                 C.POINTER(F32), P],
    "sadh_init_gradient_weighted": [P, U32, P, P, P, F32, U32, F32, F32, I32, I32, P],
    "sadh_fit_plan": [C.POINTER(FitOptions), I32, I32, C.POINTER(F32), C.POINTER(F32), C.POINTER(I32),
                      C.POINTER(I32)],
    "sadh_fit": [C.POINTER(FitOptions), P, P, I32, I32, P, I32, P, I32, C.POINTER(FitResult)],
    "sadh_fit_batch": [C.POINTER(FitOptions), C.POINTER(BatchJob), I32, C.POINTER(BatchResult)],
    "sadh_server_create": [C.POINTER(ServerOptions), C.POINTER(P)],
    "sadh_server_submit_fit": [P, C.POINTER(BatchJob), C.POINTER(C.c_longlong)],
    "sadh_server_submit_render": [P, P, U32, I32, I32, P, C.POINTER(C.c_longlong)],
    "sadh_server_step": [P, C.POINTER(I32)],
    "sadh_server_poll": [P, C.POINTER(ServerEvent), I32, C.POINTER(I32)],
    "sadh_server_destroy": [P],
    "sadh_render_sites": [P, U32, I32, I32, I32, I32, I32, I32, F32, U32, U32, U32, U32, P, C.POINTER(F32),
                          C.POINTER(F32), P],
}

_lib = None


def load() -> C.CDLL:
    """Load libsad_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SadHipError(f"{LIB_PATH} is missing: build it with `make -C sad_amd/csrc` "
                          "or __graft_entry__.build()")
    try:
        # torch ships its own HIP runtime; importing it first makes the library
        # bind to that same runtime, so torch tensors and streams can be passed in.
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.sadh_server_time.restype = C.c_double
    lib.sadh_server_time.argtypes = [P]
    lib.sadh_last_error.restype = C.c_char_p
    lib.sadh_last_error.argtypes = []
    lib.sadh_version.restype = C.c_char_p
    lib.sadh_version.argtypes = []
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call entry point `name`; raise SadHipError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.sadh_last_error().decode(errors="replace")
        raise SadHipError(f"{name} failed with status {rc}: {msg}")


def ptr(t) -> int | None:
    """Device/host address of a torch tensor or numpy array (None for None)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def declared_symbols() -> list[str]:
    """Entry points declared in include/sad_hip.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "sad_hip.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(sadh_[a-z0-9_]+)\s*\(", text)))
