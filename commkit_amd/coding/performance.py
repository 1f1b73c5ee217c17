"""Monte-Carlo error-rate curves: :func:`error_rate_curve`, :class:`ErrorRateCurve`.

``commkit/coding/performance.py:43-375``: at each Es/N0, random information
bits are encoded, mapped, sent through AWGN at one sample per symbol,
demapped and decoded in batches until ``min_block_errors`` blocks are in
error or ``max_blocks`` blocks are decoded.  Here every step runs on the
GPU through the library's kernels (encoder, Gray mapper, demapper, decoder,
and ``ck_bit_errors`` for the counts); the information bits are drawn on
the host from ``rng`` and the noise on the device from a generator seeded by
it, as in the reference.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Any

import numpy as np

from .._array import require_int
from ..mapping import Constellation
from .base import _is_code, _resolve_method, decode, encode

__all__ = ["ErrorRateCurve", "error_rate_curve"]

logger = logging.getLogger("commkit_amd")

#: Transmitted bits per batch: a few hundred long codewords, enough to fill
#: the GPU (the reference's host batches are 2^18 bits).
BATCH_BITS = 1 << 22

_COUNT_FIELDS = ("num_blocks", "bit_errors", "block_errors", "pre_fec_bit_errors",
                 "detected_failures", "undetected_errors", "total_iterations")  # fmt: skip


@dataclass(frozen=True, eq=False)
class ErrorRateCurve:
    """Decoding error counts of a code against Es/N0, one entry per point
    (``commkit/coding/performance.py:43-158``); counts are read-only int64
    host arrays, the rates properties."""

    esn0_db: np.ndarray
    num_blocks: np.ndarray
    bit_errors: np.ndarray
    block_errors: np.ndarray
    pre_fec_bit_errors: np.ndarray
    detected_failures: np.ndarray
    undetected_errors: np.ndarray
    total_iterations: np.ndarray
    k: int
    n: int
    bits_per_symbol: int
    checked: bool = True
    label: str = ""

    def __post_init__(self) -> None:
        es = np.array(self.esn0_db, dtype=np.float64)
        if es.ndim != 1 or es.size == 0 or not np.all(np.isfinite(es)):
            raise ValueError("ErrorRateCurve: esn0_db must be 1-D, finite and not empty, got "
                             f"shape {es.shape}.")  # fmt: skip
        es.setflags(write=False)
        object.__setattr__(self, "esn0_db", es)
        for name in _COUNT_FIELDS:
            v = np.asarray(getattr(self, name))
            if v.shape != es.shape:
                raise ValueError(f"ErrorRateCurve: {name} must have the shape of esn0_db "
                                 f"{es.shape}, got {v.shape}.")  # fmt: skip
            if v.dtype.kind not in "iu" or np.any(v < 0):
                raise ValueError(f"ErrorRateCurve: {name} must be counts >= 0.")
            v = v.astype(np.int64)
            v.setflags(write=False)
            object.__setattr__(self, name, v)
        if np.any(self.num_blocks < 1):
            raise ValueError("ErrorRateCurve: num_blocks must be >= 1 at every point.")
        for name in ("k", "n", "bits_per_symbol"):
            object.__setattr__(self, name, require_int(getattr(self, name),
                                                       f"ErrorRateCurve: {name}", low=1))  # fmt: skip
        object.__setattr__(self, "checked", bool(self.checked))

    @property
    def rate(self) -> float:
        return self.k / self.n

    @property
    def ebn0_db(self) -> np.ndarray:
        """Es/N0 per information bit: ``esn0_db - 10 log10(rate * bits_per_symbol)``."""
        return self.esn0_db - 10 * np.log10(self.rate * self.bits_per_symbol)

    @property
    def ber(self) -> np.ndarray:
        return self.bit_errors / (self.num_blocks * self.k)

    @property
    def fer(self) -> np.ndarray:
        return self.block_errors / self.num_blocks

    @property
    def pre_fec_ber(self) -> np.ndarray:
        return self.pre_fec_bit_errors / (self.num_blocks * self.n)

    @property
    def mean_iterations(self) -> np.ndarray:
        return self.total_iterations / self.num_blocks


def _count_rows(rx: Any, tx: Any, block: int) -> tuple[int, Any]:
    """Total bit errors and the per-block error flags (device) of two
    (rows, n) int8 tensors, from ``ck_bit_errors`` with one block a row."""
    import torch

    from .. import _native

    rows, n = int(rx.shape[0]), int(rx.shape[1])
    bit_err = torch.empty(rows, dtype=torch.int64, device=rx.device)
    blk_err = torch.empty(rows, dtype=torch.int64, device=rx.device)
    _native.check(_native.lib().ck_bit_errors(rx.data_ptr(), tx.data_ptr(), rows, n, block,
                                              bit_err.data_ptr(), blk_err.data_ptr(),
                                              _native.stream_of(rx.device)))  # fmt: skip
    return bit_err, blk_err


def error_rate_curve(code: Any, *, constellation: Constellation, esn0_db: Any,
                     method: Any = None, min_block_errors: int = 100, max_blocks: int = 10_000,
                     label: str | None = None, rng: Any = None,
                     device: str = "gpu") -> ErrorRateCurve:
    """Measure a code's error rates over AWGN by Monte Carlo
    (``commkit/coding/performance.py:172-375``).

    ``code``: LDPCCode, CCDM or PAS; ``constellation``: uniform for a FEC
    code, the code's own for a matcher; ``esn0_db`` (P,); a point stops
    after the batch that reaches ``min_block_errors`` block errors, the last
    batch cut to ``max_blocks``.  The noise is circular complex of variance
    ``N0 = Es / (Es/N0)`` (real ``N0 / 2`` for PAM), and the LLRs use that
    variance and the constellation's prior.  ``device`` is accepted for the
    reference's signature; the work runs on the current GPU either way.
    Raises TypeError / ValueError as the reference does.
    """
    import torch

    name = "error_rate_curve()"
    if not _is_code(code):
        raise TypeError(f"{name}: code must be a code object, got {type(code).__name__}.")
    method = _resolve_method(code, method, name)
    if not isinstance(constellation, Constellation):
        raise TypeError(f"{name}: constellation must be a Constellation, got "
                        f"{type(constellation).__name__}.")  # fmt: skip
    matched = getattr(code, "constellation", None)
    if matched is not None and matched != constellation:
        raise ValueError(f"{name}: {code!r} matches onto another constellation than "
                         f"{constellation!r}.")  # fmt: skip
    if constellation.pmf is not None and matched is None:
        raise ValueError(f"{name}: a shaped constellation with a FEC code needs a distribution "
                         "matcher: use code=PAS(code, constellation).")  # fmt: skip
    grid = np.array(esn0_db, dtype=np.float64)
    if grid.ndim != 1 or grid.size == 0 or not np.all(np.isfinite(grid)):
        raise ValueError(f"{name}: esn0_db must be 1-D, finite and not empty.")
    require_int(min_block_errors, f"{name}: min_block_errors", low=1)
    require_int(max_blocks, f"{name}: max_blocks", low=1)
    if device not in ("cpu", "gpu"):
        raise ValueError(f"{name}: device must be 'cpu' or 'gpu', got {device!r}.")
    if not torch.cuda.is_available():
        raise RuntimeError(f"{name}: the receive tail runs on a GPU; none is visible.")
    dev = torch.device("cuda", torch.cuda.current_device())
    gen = rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)
    noise_gen = torch.Generator(device=dev)
    noise_gen.manual_seed(int(gen.integers(0, 2**62)))
    k, n, m = int(code.k), int(code.n), constellation.bits_per_symbol
    es = constellation.power()
    batch = max(1, -(-BATCH_BITS // n))
    if label is None:
        label = f"{type(method).__name__ if method is not None else 'Dematcher'} on {code!r}"
    counts = np.zeros((len(_COUNT_FIELDS), grid.size), dtype=np.int64)
    for i, esn0 in enumerate(grid):
        n0 = es / 10 ** (esn0 / 10)
        noise_var = n0 if constellation.is_complex else n0 / 2
        while counts[0, i] < max_blocks and counts[2, i] < min_block_errors:
            b = int(min(batch, max_blocks - counts[0, i]))
            info = torch.from_numpy(gen.integers(0, 2, (b, k), dtype=np.int8)).to(dev)
            bits = encode(info, code=code).reshape(-1)
            filler = -bits.numel() % m
            if filler:
                extra = torch.from_numpy(gen.integers(0, 2, filler, dtype=np.int8)).to(dev)
                bits = torch.cat([bits, extra])
            sym = constellation.map(bits)
            if sym.is_complex():
                std = float(np.sqrt(noise_var / 2))
                noise = torch.complex(
                    torch.randn(sym.shape, generator=noise_gen, device=dev),
                    torch.randn(sym.shape, generator=noise_gen, device=dev)) * std  # fmt: skip
            else:
                noise = torch.randn(sym.shape, generator=noise_gen, device=dev) * float(
                    np.sqrt(noise_var))
            llrs = constellation.llr((sym + noise).to(sym.dtype), noise_var=float(noise_var))
            llrs = llrs.reshape(-1)[: b * n].reshape(b, n)
            res = decode(llrs, code=code, method=method)
            bit_err, blk_err = _count_rows(res.bits.contiguous(), info, k)
            hard = (llrs < 0).to(torch.int8)
            pre, _ = _count_rows(hard, bits[: b * n].reshape(b, n).contiguous(), n)
            valid = res.valid.reshape(-1)
            wrong = blk_err > 0
            batch_counts = torch.stack([
                torch.tensor(b, dtype=torch.int64, device=dev),
                bit_err.sum(), wrong.sum(), pre.sum(), (~valid).sum(),
                (wrong & valid).sum(), res.iterations.to(torch.int64).sum()])  # fmt: skip
            counts[:, i] += batch_counts.cpu().numpy()
        logger.debug("  Es/N0 %.2f dB: %d block errors in %d blocks, %d bit errors.", esn0,
                     counts[2, i], counts[0, i], counts[1, i])  # fmt: skip
    return ErrorRateCurve(grid, *counts, k=k, n=n, bits_per_symbol=m, checked=True,
                          label=label)  # fmt: skip
