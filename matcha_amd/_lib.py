"""ctypes binding of libmatcha_amd.so (the C-ABI of ``include/matcha_amd.h``).

The product path has no CPU fallback: if the shared library is missing or a
device is not available, every entry point raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libmatcha_amd.so"

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_float = ctypes.c_float
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
c_int_p = ctypes.POINTER(ctypes.c_int)


# (name, restype, argtypes)
_SIGNATURES = [
    ("mch_last_error", ctypes.c_char_p, []),
    ("mch_so3_plan_create", c_int, [c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("mch_so3_plan_destroy", c_int, [c_void_p]),
    ("mch_sofft_topk", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p]),
    ("mch_sofft_eval", c_int, [c_void_p, c_void_p, c_ll, c_int, c_void_p, c_void_p]),
    ("mch_newton_march", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_int_p, c_int, c_int,
                                 c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mch_so3_set_stop_early", c_int, [c_void_p, c_int]),
    ("mch_reduce_candidates", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p]),
    ("mch_search_orientations", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_int_p, c_int, c_int,
                                        c_int, c_float, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    ("mch_search_orientations_ends", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_int_p, c_int,
                                             c_void_p, c_int, c_int, c_float, c_void_p, c_void_p,
                                             c_void_p, c_void_p, c_void_p]),
    ("mch_so3_evaluate", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_int, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    ("mch_fle_plan_create", c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_int, c_int, ctypes.POINTER(c_void_p)]),
    ("mch_fle_plan_destroy", c_int, [c_void_p]),
    ("mch_fle_build_basis", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_double]),
    ("mch_fle_copy_basis", c_int, [c_void_p, c_void_p, c_void_p]),
    ("mch_fle_project", c_int, [c_void_p, c_void_p, c_ll, c_int, c_void_p, c_ll, c_void_p]),
    ("mch_fle_project_shifted", c_int, [c_void_p, c_void_p, c_ll, c_int, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_fle_project_tangents", c_int, [c_void_p, c_void_p, c_ll, c_int, c_int, c_void_p, c_void_p, c_ll,
                                          c_void_p]),
    ("mch_fle_gemm_only", c_int, [c_void_p, c_int, c_void_p]),
    ("mch_fle_set_precision", c_int, [c_void_p, c_int]),
    ("mch_fle_tuned_depth", c_int, [c_void_p]),
    ("mch_fle_set_gemm_tiles", c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    ("mch_corr_plan_create", c_int, [c_int, c_int, c_void_p, c_int, ctypes.POINTER(c_void_p)]),
    ("mch_corr_plan_destroy", c_int, [c_void_p]),
    ("mch_compute_sigma", c_int, [c_void_p, c_void_p, c_void_p, c_ll, c_int, c_int, c_void_p, c_ll,
                                  c_void_p, c_void_p]),
    ("mch_compute_sigma_degree", c_int, [c_void_p, c_void_p, c_void_p, c_ll, c_int, c_int, c_int, c_void_p,
                                         c_ll, c_void_p, c_void_p]),
    ("mch_shift_plan_create", c_int, [c_int, c_void_p, c_float, c_int, ctypes.POINTER(c_void_p)]),
    ("mch_shift_plan_destroy", c_int, [c_void_p]),
    ("mch_shift_plan_set_upsample", c_int, [c_void_p, c_int, c_int]),
    ("mch_shift_search", c_int, [c_void_p, c_void_p, c_ll, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    ("mch_shift_plan_set_timing", c_int, [c_void_p, c_int]),
    ("mch_shift_set_precision", c_int, [c_void_p, c_int]),
    ("mch_shift_kernel_ms", c_int, [c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ("mch_extract_patches", c_int, [c_void_p, c_ll, c_int, c_int, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_euler_to_shift_rot", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    ("mch_extract_rotated", c_int, [c_void_p, c_ll, c_int, c_int, c_int, c_void_p, c_void_p, c_double,
                                    c_void_p, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_tracker_update", c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    ("mch_score_pose", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("mch_shift_tangent", c_int, [c_void_p, c_ll, c_int, c_int, c_int, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_score_gradients", c_int, [c_void_p, c_int, c_int, c_void_p, c_ll, c_int, c_void_p, c_ll, c_int,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_score_gradients_noalloc", c_int, [c_void_p, c_int, c_int, c_void_p, c_ll, c_int, c_void_p, c_ll,
                                            c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_ll, c_void_p]),
    ("mch_score_backward", c_int, [c_void_p, c_void_p, c_int, c_void_p, c_ll, c_int, c_void_p, c_void_p,
                                   c_void_p, c_void_p]),
    ("mch_wigner_d_host", c_int, [c_double, c_int, ctypes.POINTER(c_double)]),
]


def lib():
    """Load the library once; raise loudly if it was not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m matcha_amd.build` "
                "(the HIP path has no fallback)")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, res, args in _SIGNATURES:
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def exported_symbols():
    return [name for name, _, _ in _SIGNATURES]


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().mch_last_error()
        msg = msg.decode(errors="replace") if msg else ""
        raise RuntimeError(f"matcha_amd {what} failed (code {rc}): {msg}")


def stream_handle(device: torch.device) -> int:
    """Raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu_tensor(t: torch.Tensor, name: str, dtype=None, device=None) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on a GPU device (got {t.device}); "
                         "matcha_amd has no CPU path")
    if device is not None and t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    if dtype is not None and t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()
