"""The coding verbs :func:`encode` and :func:`decode`, and :class:`DecodeResult`.

Contract of ``commkit/coding/base.py:46-379``: LLRs of the transmitted bits
are ``(N,)`` or ``(C, N)`` with ``N`` a multiple of ``code.n``, blocks
concatenated along the last axis; float32 (float64 warns and is reduced),
integer or bool input is read as hard decisions (LLR ``+-1``).  The result
arrays are on the input's device: tensors for device input, NumPy for host
input (host input is copied to the GPU, decoded by the HIP kernels and
copied back).
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Any

import numpy as np

from .._array import as_2d, restore_1d, warn_if_double
from ..backend import dispatch, to_host_like
from . import ldpc as _ldpc
from . import shaping as _shaping

__all__ = ["DecodeResult", "decode", "encode"]

logger = logging.getLogger("commkit_amd")

# Magnitude of the LLR given to an integer hard decision.
_HARD_LLR = 1.0


@dataclass(frozen=True, eq=False)
class DecodeResult:
    """Output of :func:`decode`, arrays on the input's device.

    bits : int8 ``(..., B * k)`` decoded information bits.
    llrs : float32 ``(..., B * n)`` a-posteriori LLRs of the transmitted bits.
    valid : bool ``(..., B)``: the syndrome check passed (detected success).
    iterations : int32 ``(..., B)`` iterations run.
    checked : whether ``valid`` comes from a check that can fail (True).
    """

    bits: Any
    llrs: Any | None
    valid: Any
    iterations: Any
    checked: bool


_CODE_CLASSES = "LDPCCode, CCDM, PAS"


def _is_code(code: Any) -> bool:
    return isinstance(code, (_ldpc.LDPCCode, _shaping.CCDM, _shaping.PAS))


def _resolve_method(code: Any, method: Any, function_name: str) -> Any:
    if not _is_code(code):
        raise TypeError(
            f"{function_name}: code must be a code object ({_CODE_CLASSES}), "
            f"got {type(code).__name__}."
        )
    if isinstance(code, _shaping.CCDM):
        # The matcher's only decoder inverts the hard decisions
        # (commkit/coding/shaping.py:223-246).
        if method is not None:
            raise TypeError(
                f"{function_name}: method {type(method).__name__} does not decode "
                "CCDM; its only decoder is the dematcher (method=None)."
            )
        return None
    if isinstance(code, _shaping.PAS):
        return _resolve_method(code.code, method, function_name)
    if method is None:
        return _ldpc.MinSum()
    if not isinstance(method, (_ldpc.MinSum, _ldpc.SumProduct)):
        raise TypeError(
            f"{function_name}: method {type(method).__name__} does not decode "
            f"{type(code).__name__}; use one of MinSum, SumProduct."
        )
    return method


def _blocks(x: Any, size: int, name: str, function_name: str) -> Any:
    if x.shape[-1] % size:
        raise ValueError(
            f"{function_name}: the last axis of {name} ({x.shape[-1]}) must be a "
            f"multiple of the block length {size}."
        )
    return x.reshape(-1, size)


def _as_llrs(t: Any, function_name: str) -> Any:
    import torch

    if t.is_complex():
        raise TypeError(f"{function_name}: llrs must be real, got {t.dtype}.")
    if t.dtype == torch.bool or not t.dtype.is_floating_point:
        hard = (t != 0).to(torch.float32)
        return hard * (-2.0 * _HARD_LLR) + _HARD_LLR
    warn_if_double(t.dtype, function_name, "float32")
    return t.to(torch.float32)


def decode(llrs: Any, *, code: Any, method: Any = None) -> DecodeResult:
    """Decode LLRs (or hard decisions) of a code's transmitted bits.

    Raises TypeError for a non-code or a method that does not decode it,
    ValueError when the last axis is not a multiple of ``code.n``.
    """
    method = _resolve_method(code, method, "decode()")
    placed = dispatch(llrs, "llrs")
    x = _as_llrs(placed.tensor, "decode()")
    if x.ndim == 0:
        raise ValueError("llrs must be 1-D (N,) or 2-D (C, N), got 0-D.")
    x2, was_1d = as_2d(x, name="llrs")
    blocks = _blocks(x2, code.n, "llrs", "decode()")
    logger.debug("Decoding %d block(s) of %r with %r.", blocks.shape[0], code, method)
    if isinstance(code, _shaping.CCDM):
        # Hard decisions of the labels, then the dematcher; no soft output.
        import torch

        hard = (blocks < 0).to(torch.int8)
        bits, valid = code._dematch_labels(hard)
        app = None
        iters = torch.ones(blocks.shape[0], dtype=torch.int32, device=blocks.device)
    elif isinstance(code, _shaping.PAS):
        bits, app, valid, iters = code._decode(blocks, method)
    else:
        bits, app, valid, iters = _ldpc._decode(code, blocks, method)
    num_blocks = x2.shape[-1] // code.n
    rows = x2.shape[0]

    def shaped(a: Any, size: int) -> Any:
        return restore_1d(was_1d, a.reshape(rows, num_blocks * size))

    if logger.isEnabledFor(logging.INFO) and iters.numel():
        it_host = iters.cpu().numpy()
        failed = int(valid.numel() - int(valid.sum()))
        logger.info("Decoded %d block(s) of %r with %r: %d failed the decoder's check, "
                    "iterations mean %.1f, max %d.", it_host.size, code, method, failed,
                    float(it_host.mean()), int(it_host.max()))  # fmt: skip
    kind = placed.kind
    return DecodeResult(
        bits=to_host_like(shaped(bits, code.k), kind),
        llrs=None if app is None else to_host_like(shaped(app, code.n), kind),
        valid=to_host_like(shaped(valid, 1), kind),
        iterations=to_host_like(shaped(iters, 1), kind),
        checked=True,
    )


def encode(bits: Any, *, code: Any) -> Any:
    """Encode information bits ``(N,)`` or ``(C, N)``, ``N`` a multiple of ``code.k``,
    into int8 transmitted bits ``(..., N / k * n)`` on the input's device."""
    import torch

    if not _is_code(code):
        raise TypeError(
            f"encode(): code must be a code object ({_CODE_CLASSES}), got "
            f"{type(code).__name__}."
        )
    placed = dispatch(bits, "bits")
    x = placed.tensor
    if x.is_complex() or x.dtype.is_floating_point:
        raise TypeError(f"encode(): bits must be integers or bool, got {x.dtype}.")
    if x.dtype != torch.bool and x.numel() and bool(((x < 0) | (x > 1)).any()):
        raise ValueError("encode(): bits must be 0 or 1.")
    x2, was_1d = as_2d(x, name="bits")
    blocks = _blocks(x2, code.k, "bits", "encode()").to(torch.int8)
    if blocks.shape[0] == 0:
        out = torch.empty((0, code.n), dtype=torch.int8, device=x.device)
    elif isinstance(code, (_shaping.CCDM, _shaping.PAS)):
        out = code._encode(blocks)
    else:
        out = _ldpc._encode(code, blocks)
    num_blocks = x2.shape[-1] // code.k
    res = restore_1d(was_1d, out.reshape(x2.shape[0], num_blocks * code.n))
    return to_host_like(res, placed.kind)


def hard_decisions(llrs: Any) -> Any:
    """int8 1 where the LLR is negative (``commkit/coding/base.py:227-229``)."""
    return (np.asarray(llrs) < 0).astype(np.int8)
