"""Random payloads on the device: :func:`generate`.

``commkit/core/generation.py:167-405`` at one sample per symbol (pulse
shaping is outside the receive tail): information bits are drawn on the
host from ``rng`` (so a seed gives the same bits on every device), encoded
by the code's encoder kernel (LDPC, CCDM, PAS), mapped by the
constellation, and returned as a Signal whose reference holds the symbols,
the bits on the air, the information bits and the code.
"""

from __future__ import annotations

import logging
import math
from typing import Any

import numpy as np

from .._array import require_int
from .signal import Reference, Signal

__all__ = ["generate"]

logger = logging.getLogger("commkit_amd")


def _coded_layout(num_symbols: int | None, num_blocks: int | None, m: int, n: int,
                  pad: bool) -> tuple[int, int]:
    """(symbols, blocks) of a record of ``n``-bit blocks on ``m``-bit symbols
    (``commkit/core/generation.py:408-446``)."""
    step_symbols = n // math.gcd(n, m)
    step_blocks = m // math.gcd(n, m)
    if num_blocks is not None:
        bits = num_blocks * n
        if bits % m and not pad:
            low = num_blocks // step_blocks * step_blocks
            valid = f"{low} or {low + step_blocks}" if low else f"{step_blocks}"
            raise ValueError(f"generate(): num_blocks={num_blocks} fills {bits / m:g} symbols of "
                             f"{m} bits; use num_blocks={valid}, or pad=True to fill the last "
                             "symbol with random bits.")  # fmt: skip
        return -(-bits // m), num_blocks
    assert num_symbols is not None
    total = num_symbols * m
    blocks = total // n
    if blocks == 0:
        raise ValueError(f"generate(): num_symbols={num_symbols} carries {total} bits, fewer "
                         f"than one code block of {n}.")  # fmt: skip
    if total % n and not pad:
        low = num_symbols // step_symbols * step_symbols
        valid = f"{low} or {low + step_symbols}" if low else f"{step_symbols}"
        raise ValueError(f"generate(): num_symbols={num_symbols} carries {total} bits, not a "
                         f"multiple of the code's block length {n}; use num_symbols={valid}, "
                         "num_blocks= to size the record in blocks, or pad=True to fill the "
                         "rest with random bits.")  # fmt: skip
    return num_symbols, blocks


def generate(constellation: Any, *, num_symbols: int | None = None, symbol_rate: float,
             sps: int = 1, pulse: Any = None, num_channels: int = 1, periodic: bool = False,
             code: Any = None, num_blocks: int | None = None, pad: bool = False,
             rng: Any = None, device: str = "cpu") -> Signal:
    """Random symbols of ``constellation`` (coded by ``code``) as a Signal at
    one sample per symbol (``commkit/core/generation.py:167-405``).

    A shaped constellation without a code goes through one
    ``CCDM(constellation, num_symbols)`` block per channel.  ``device``
    "cpu" returns NumPy arrays, "gpu" tensors on the current GPU; the work
    runs on the GPU either way.  Raises ValueError as the reference does,
    and for ``sps != 1`` or a pulse.
    """
    import torch

    from .. import coding
    from ..mapping import Constellation

    name = "generate()"
    if not isinstance(constellation, Constellation):
        raise TypeError(f"{name}: constellation must be a Constellation, e.g. "
                        f"Constellation.qam(16); got {type(constellation).__name__}.")  # fmt: skip
    if (num_symbols is None) == (num_blocks is None):
        raise ValueError(f"{name}: give exactly one of num_symbols and num_blocks.")
    if code is None and (num_blocks is not None or pad):
        raise ValueError(f"{name}: num_blocks and pad size a coded record; they need code=.")
    for arg, value in (("num_symbols", num_symbols), ("num_blocks", num_blocks),
                       ("num_channels", num_channels)):  # fmt: skip
        if value is not None and require_int(value, f"{name}: {arg}") < 1:
            raise ValueError(f"{name}: {arg} must be >= 1, got {value}.")
    if sps != 1 or pulse is not None or periodic:
        raise ValueError(f"{name}: pulse shaping is outside the receive tail; generate "
                         "symbols at sps=1 without a pulse.")  # fmt: skip
    if not (isinstance(symbol_rate, (int, float)) and symbol_rate > 0):
        raise ValueError(f"{name}: symbol_rate must be positive, got {symbol_rate!r}.")
    if device not in ("cpu", "gpu"):
        raise ValueError(f"{name}: device must be 'cpu' or 'gpu', got {device!r}.")
    if not torch.cuda.is_available():
        raise RuntimeError(f"{name}: the encoder and mapper run on a GPU; none is visible.")
    dev = torch.device("cuda", torch.cuda.current_device())
    gen = rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)
    c = constellation
    k = c.bits_per_symbol
    info = None
    if code is not None or c.pmf is not None:
        if code is None:
            code = coding.CCDM(c, num_symbols=int(num_symbols))
        matched = getattr(code, "constellation", None)
        if matched is not None and matched != c:
            raise ValueError(f"{name}: {code!r} matches onto another constellation than {c!r}.")
        if c.pmf is not None and matched is None:
            raise ValueError(f"{name}: a shaped constellation with a FEC code needs a "
                             "distribution matcher: use code=PAS(code, constellation).")  # fmt: skip
        if pad and matched is not None:
            raise ValueError(f"{name}: pad fills with uniform random bits, which would change "
                             f"the composition of {code!r}; size the record in whole blocks.")  # fmt: skip
        num_symbols, num_blocks = _coded_layout(num_symbols, num_blocks, k, code.n, pad)
        info_h = gen.integers(0, 2, size=(num_channels, num_blocks * code.k), dtype=np.int8)
        info = torch.from_numpy(info_h).to(dev)
        bits = coding.encode(info, code=code)
        filler = num_symbols * k - num_blocks * code.n
        if filler:
            extra = gen.integers(0, 2, size=(num_channels, filler), dtype=np.int8)
            bits = torch.cat([bits, torch.from_numpy(extra).to(dev)], dim=1)
    else:
        bits_h = gen.integers(0, 2, size=(num_channels, int(num_symbols) * k), dtype=np.int8)
        bits = torch.from_numpy(bits_h).to(dev)
    symbols = c.map(bits)
    if num_channels == 1:
        bits, symbols = bits[0], symbols[0]
        info = None if info is None else info[0]

    def out(t: Any) -> Any:
        return None if t is None else (t.cpu().numpy() if device == "cpu" else t)

    logger.info("Generated %r%s: %s symbols x %s channel(s) on %s.", c,
                "" if code is None else f" coded with {code!r}", num_symbols, num_channels,
                device)  # fmt: skip
    sym = out(symbols)
    return Signal(samples=sym, constellation=c, sps=1, symbol_rate=float(symbol_rate),
                  reference=Reference(symbols=sym, bits=out(bits), info_bits=out(info),
                                      code=code))  # fmt: skip
