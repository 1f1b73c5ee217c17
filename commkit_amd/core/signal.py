"""The :class:`Signal` record, as far as the receive tail reads it.

``compute_llr`` takes an array or a Signal (``commkit/core/_signal_input.py:
46,216-240,309-327``): a Signal must be at one sample per symbol and carry
no frame, and supplies the constellation when the call does not.  Only the
fields that contract reads are kept here: the samples, the constellation,
samples per symbol and the frame, plus the reference of what was sent,
which scoring (:mod:`commkit_amd.metrics`) and storage (:mod:`commkit_amd.io`)
read.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

__all__ = ["Reference", "Signal", "signal_input"]


@dataclass(frozen=True, eq=False)
class Reference:
    """What was sent (``commkit/core/signal.py``'s Reference, as far as scoring
    and storage read it).

    symbols : the transmitted symbols.
    bits : the bits on the air (the labels of ``symbols``).
    info_bits : the source bits before ``code`` (``bits`` when uncoded).
    code : the code between them (LDPCCode, CCDM, PAS) or ``None``.
    """

    symbols: Any = None
    bits: Any = None
    info_bits: Any = None
    code: Any = None


@dataclass(frozen=True, eq=False)
class Signal:
    """Received samples with their constellation and sampling.

    samples : array ``(..., N)`` (NumPy or a device tensor).
    constellation : the transmitted :class:`~commkit_amd.mapping.Constellation`.
    sps : samples per symbol (the tail needs 1).
    frame : framing of the record, ``None`` when the samples are payload only.
    reference : the :class:`Reference` of what was sent, or ``None``.
    symbol_rate : symbols per second, or ``None`` when not recorded.
    """

    samples: Any
    constellation: Any = None
    sps: int = 1
    frame: Any = None
    reference: Any = None
    symbol_rate: float | None = None


def signal_input(symbols: Any, constellation: Any, function_name: str) -> tuple[Any, Any]:
    """``(samples, constellation)`` from an array or a Signal, checked."""
    if isinstance(symbols, Signal):
        if symbols.sps != 1:
            raise ValueError(
                f"{function_name} needs symbols at 1 sample per symbol; the Signal "
                f"has sps={symbols.sps}. Downsample it first."
            )
        if symbols.frame is not None:
            raise ValueError(
                f"{function_name} takes payload symbols; the Signal carries a frame. "
                "Extract the payload first."
            )
        c = constellation if constellation is not None else symbols.constellation
        if c is None:
            raise ValueError(f"{function_name}: the Signal has no constellation; pass one.")
        return symbols.samples, c
    if constellation is None:
        raise ValueError(f"{function_name}: constellation is required for array input.")
    return symbols, constellation
