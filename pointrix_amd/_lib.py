"""ctypes binding of libgsr.so (declared in include/gsr.h).

The library is loaded from this package directory only; there is no fallback:
calling an op without the built library raises immediately.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsr.so")

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); the single source of truth for the Python side
SIGNATURES = {
    "gsr_version": (_I, []),
    "gsr_last_error": (ctypes.c_char_p, []),
    "gsr_project_point_fwd": (_I, [_P, _I64, _P, _P, _I, _I, _F, _F, _P, _P, _P]),
    "gsr_project_point_bwd": (_I, [_P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gsr_compute_cov3d_fwd": (_I, [_P, _P, _P, _I64, _P, _P]),
    "gsr_compute_cov3d_bwd": (_I, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "gsr_ewa_project_fwd": (_I, [_P, _P, _I64, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P]),
    "gsr_ewa_project_bwd": (_I, [_P, _P, _I64, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gsr_compute_sh_fwd": (_I, [_P, _I64, _I64, _I64, _P, _I64, _I, _I, _I, _P, _P]),
    "gsr_compute_sh_bwd": (_I, [_P, _I64, _I64, _I64, _P, _I64, _I, _I, _I, _P, _P, _P, _P]),
    "gsr_preprocess_fwd": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _P, _P, _P, _I, _I, _F, _F, _I,
                                _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gsr_preprocess_bwd": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _P, _P, _P, _I, _I, _F, _F, _I,
                                _P, _P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                                _P, _P, _I, _P, _P]),
    "gsr_sort_ws_gauss_bytes": (_SZ, [_I64]),
    "gsr_sort_ws_inst_bytes": (_SZ, [_I64]),
    "gsr_sort_count": (_I, [_P, _P, _P, _I64, _I, _I, _I, _I, _P, _P, _P]),
    "gsr_sort_emit": (_I, [_P, _P, _I64, _I, _I, _I, _I, _P, _I64, _P, _P, _P, _P]),
    "gsr_set_binning_mode": (_I, [_I]),
    "gsr_sort_gaussian_noalloc": (_I, [_P, _P, _P, _I64, _I, _I, _I, _I, _I64, _P, _P, _P, _P]),
    "gsr_mv_preprocess_fwd": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _I, _P, _P, _I, _I, _F, _F,
                                   _I, _P, _P, _P, _P, _P, _P, _P]),
    "gsr_mv_preprocess_bwd": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _I, _P, _P, _F, _F, _I, _P,
                                   _P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gsr_sort_count_tight": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _I, _I, _P, _P, _P]),
    "gsr_sort_count_high": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _I, _I, _P, _P, _P, _P]),
    "gsr_sort_count_async": (_I, [_P, _P, _P, _P, _P, _I64, _I, _I, _I, _I, _P, _P, _P, _P,
                                  _P]),
    "gsr_sort_emit_tight": (_I, [_P, _P, _P, _P, _I64, _I, _I, _I, _I, _P, _I64, _P, _P, _P,
                                 _P]),
    "gsr_mv_sort_count": (_I, [_P, _P, _P, _P, _P, _I64, _I, _P, _I, _I, _P, _P, _P]),
    "gsr_mv_sort_count_async": (_I, [_P, _P, _P, _P, _P, _I64, _I, _P, _I, _I, _P, _P, _P, _P,
                                     _P]),
    "gsr_mv_sort_emit": (_I, [_P, _P, _P, _P, _I64, _I, _P, _I, _I, _I, _P, _I64, _P, _P, _P,
                              _P]),
    "gsr_mv_alpha_blending_fwd": (_I, [_P, _P, _P, _P, _I, _P, _P, _F, _I64, _I, _P, _I, _I64, _I,
                                       _I, _I, _P, _P, _P, _P, _P]),
    "gsr_mv_alpha_blending_bwd": (_I, [_P, _P, _P, _P, _I, _P, _P, _F, _I64, _I, _P, _I, _I64, _I,
                                       _I, _I, _P, _P, _P, _P, _P, _P, _I, _P]),
    "gsr_ssim_partials": (_I, [_I, _I, _I]),
    "gsr_l1_ssim_fwd": (_I, [_P, _P, _I, _I, _I, _P, _P, _P]),
    "gsr_l1_ssim_bwd": (_I, [_P, _P, _I, _I, _I, _P, _P, _P]),
    "gsr_adam_chunk": (_I, []),
    "gsr_adam_entry_bytes": (_I, []),
    "gsr_adam_step": (_I, [_P, _I, _I64, _D, _D, _D, _I, _P]),
    "gsr_row_copy_bytes": (_I, []),
    "gsr_gather_rows": (_I, [_P, _I, _P, _I64, _P]),
    "gsr_densify_stats": (_I, [_P, _I64, _I64, _I, _I64, _P, _P, _F, _F, _P, _P, _P, _P]),
    "gsr_uv_grad_scale": (_I, [_P, _I, _I64, _F, _F, _P, _P]),
    "gsr_activate_fwd": (_I, [_P, _P, _P, _I64, _P, _P, _P, _P]),
    "gsr_activate_bwd": (_I, [_P, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P]),
    "gsr_camera_pose_fwd": (_I, [_P, _P, _I, _P, _P, _P]),
    "gsr_camera_pose_bwd": (_I, [_P, _P, _I, _P, _P, _P, _P, _P]),
    "gsr_plane_grads_bytes": (_I, []),
    "gsr_planes_fwd": (_I, [_P, _I, _I64, _I, _P, _P]),
    "gsr_planes_bwd": (_I, [_P, _I, _I64, _I, _P, _P, _P]),
    "gsr_l1_loss_partials": (_I, [_I64]),
    "gsr_l1_loss_fwd": (_I, [_P, _P, _I64, _P, _P, _P]),
    "gsr_l1_loss_bwd": (_I, [_P, _P, _I64, _P, _P, _P]),
    "gsr_l1_terms_bytes": (_I, []),
    "gsr_l1_sum_partials": (_I, [_P]),
    "gsr_l1_sum_fwd": (_I, [_P, _P, _P, _P]),
    "gsr_l1_sum_bwd": (_I, [_P, _P, _P]),
    "gsr_alpha_blending_fwd": (_I, [_P, _P, _P, _P, _I, _P, _P, _F, _I, _I, _I, _I, _P, _P, _P,
                                    _P]),
    "gsr_alpha_blending_bwd": (_I, [_P, _P, _P, _P, _I, _P, _P, _F, _I, _I, _I, _I, _P, _P, _P,
                                    _P, _P, _I, _P]),
    "gsr_alpha_blending_bwd_split": (_I, [_P, _P, _P, _P, _I, _P, _P, _F, _I, _I, _I, _I, _P,
                                          _P, _P, _P, _P, _I, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libgsr.so (once) and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise RuntimeError(
                    f"{path} is not built; run `python -m pointrix_amd.build` first")
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


class GsrError(RuntimeError):
    pass


def call(name: str, *args):
    """Call a status-returning entry point; raise GsrError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.gsr_last_error().decode(errors="replace")
        raise GsrError(f"{name} failed with status {rc}: {msg}")
    return rc


def ptr(t: torch.Tensor | None):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(t: torch.Tensor):
    """hipStream_t of the current torch stream on t's device."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)
