"""ctypes binding of ``libck_hip.so``, the HIP library of the receive tail.

The library is built in-tree for gfx950 by :mod:`commkit_amd.build` and
exports the C ABI declared in ``include/ck_tail.h``.  There is no fallback:
if the library is missing or fails to load, every call raises, as the
reference's kernel registry does (``commkit/_cuda/__init__.py:1-9``).  torch
is imported before the library is opened so that the library binds to the
same HIP runtime as torch (both carry the soname ``libamdhip64.so.7``).

Every entry point returns 0 or a nonzero status; :func:`check` turns a
nonzero status into ``RuntimeError(ck_last_error())``.
"""

from __future__ import annotations

import ctypes
import os
from typing import Any

__all__ = ["LIB_PATH", "SYMBOLS", "lib", "check", "stream_of"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libck_hip.so")

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double

#: Exported symbol -> (restype, argtypes); mirrors include/ck_tail.h.
SYMBOLS: dict[str, tuple[Any, list[Any]]] = {
    "ck_version": (_i32, []),
    "ck_last_error": (ctypes.c_char_p, []),
    "ck_llr_separable_maxlog": (_i32, [_p, _i64, _p, _p, _f64, _p, _p]),
    "ck_llr_separable": (_i32, [_p, _i64, _p, _p, _f64, _i32, _p, _p]),
    "ck_llr_separable_rows": (_i32, [_p, _i64, _i64, _p, _p, _p, _i32, _i32, _p, _p]),
    "ck_llr_separable_segments": (_i32, [_p, _i32, _p, _p, _i32, _i32, _p]),
    "ck_llr_points": (_i32, [_p, _i64, _p, _p, _p, _i32, _i32, _f64, _i32, _p, _p]),
    "ck_llr_separable_permuted": (_i32, [_p, _i64, _p, _p, _f64, _i32, _i32, _p, _p, _i64, _i64,
                                         _p]),
    "ck_ldpc_workspace_bytes": (_i64, [_p, _p, _i64]),
    "ck_ldpc_decode": (_i32, [_p, _p, _p, _p, _p, _p, _p, _i64, _p, _p]),
    "ck_ldpc_decode_static": (_i32, [_p, _p, _p, _p, _p, _p, _p, _i64, _p]),
    "ck_ldpc_decode_symbols": (_i32, [_p, _p, _p, _p, _p, _f64, _p, _p, _p, _p, _p, _i64, _p, _p]),
    "ck_batch_workspace_bytes": (_i64, [_p, _i32, _p]),
    "ck_batch_prepare": (_i32, [_p, _i32, _p, _p, _i64, _p]),
    "ck_batch_launch": (_i32, [_p, _p]),
    "ck_batch_graph_create": (_i32, [_p, _p]),
    "ck_batch_graph_launch": (_i32, [_p, _p]),
    "ck_batch_graph_destroy": (_i32, [_p]),
    "ck_serve_workspace_bytes": (_i64, [_p]),
    "ck_serve_create": (_i32, [_p, _p, _p, _p]),
    "ck_serve_admit": (_i32, [_p, _i32, _i64, ctypes.c_uint64, _p]),
    "ck_serve_submit": (_i32, [_p, ctypes.c_uint64]),
    "ck_serve_step": (_i32, [_p, _p]),
    "ck_serve_poll": (_i32, [_p, _p, _i32, _p]),
    "ck_serve_release": (_i32, [_p, ctypes.c_uint64]),
    "ck_serve_usage": (_i32, [_p, _i32, _p, _p, _p]),
    "ck_serve_destroy": (_i32, [_p]),
    "ck_ccdm_match": (_i32, [_p, _i64, _i64, _p, _i32, _i64, _p, _p]),
    "ck_ccdm_dematch_workspace_bytes": (_i64, [_i64, _i64]),
    "ck_ccdm_dematch": (_i32, [_p, _i64, _i64, _p, _i32, _i64, _i64, _p, _p, _p, _p]),
    "ck_bit_errors": (_i32, [_p, _p, _i64, _i64, _i64, _p, _p, _p]),
    "ck_gmi_loss": (_i32, [_p, _p, _i64, _i64, _p, _p]),
    "ck_ldpc_encode": (_i32, [_p, _p, _i64, _p, _p]),
    "ck_awgn": (_i32, [_p, _i64, _i64, _i32, _p, ctypes.c_uint64, ctypes.c_uint64, _p, _p]),
}

_LIB: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises RuntimeError when it is not built."""
    global _LIB
    if _LIB is None:
        import torch  # noqa: F401  (binds the HIP runtime torch carries)

        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m commkit_amd.build` "
                "(hipcc, gfx950).  The receive tail has no CPU fallback."
            )
        handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in SYMBOLS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(status: int) -> None:
    if status != 0:
        msg = lib().ck_last_error()
        raise RuntimeError((msg or b"unknown error").decode())


def stream_of(device: Any) -> int:
    """The raw ``hipStream_t`` of torch's current stream on ``device``."""
    import torch

    return int(torch.cuda.current_stream(device).cuda_stream)
