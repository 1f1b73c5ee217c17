"""Soft-decision demapping: :func:`compute_llr`.

Same signature, arguments, errors and output layout as
``commkit/mapping/llr.py:24-95``: max-log or exact (log-sum-exp) LLRs with
the constellation's prior, float32 ``(..., N * k)`` with the ``k`` bits of a
symbol adjacent, MSB first; a positive LLR means bit 0.  Computed by the HIP
demapper on the GPU holding the data (host input goes there and back).
"""

from __future__ import annotations

import logging
from typing import Any

from ..backend import dispatch, to_host_like
from ..core.signal import signal_input
from .constellation import Constellation
from .demap import check_method, llr_any, noise_variance

__all__ = ["compute_llr"]


def _is_complex(x: Any) -> bool:
    is_complex = getattr(x, "is_complex", None)
    if callable(is_complex):
        return bool(is_complex())
    import numpy as np

    return bool(np.iscomplexobj(x))

logger = logging.getLogger("commkit_amd")


def compute_llr(
    symbols: Any,
    *,
    noise_var: float | None = None,
    esn0_db: float | None = None,
    constellation: Constellation | None = None,
    method: str = "maxlog",
) -> Any:
    """Bit LLRs for soft-decision decoding.

    ``symbols`` is ``(..., N)`` at one sample per symbol on the
    constellation's scale, or a :class:`~commkit_amd.core.signal.Signal` at
    1 sps without a frame.  Give one of ``noise_var`` (``E[|n|^2]``) and
    ``esn0_db`` (``N0 = Es / 10^(EsN0/10)``, halved for real symbols).
    ``constellation`` defaults to the Signal's.  ``method`` is ``"maxlog"``
    or ``"exact"``.

    Raises ValueError when neither or both noise arguments are given, for an
    unknown method, an oversampled or framed Signal, or more than 16 bits
    per symbol.
    """
    x, constellation = signal_input(symbols, constellation, "compute_llr()")
    check_method(method)
    nv = noise_variance(noise_var, esn0_db, constellation, not _is_complex(x), "compute_llr()")
    placed = dispatch(x, "symbols")
    t = placed.tensor
    logger.debug("Computing LLRs for %r (method=%s).", constellation, method)
    from .._array import warn_if_double

    warn_if_double(t.dtype, "compute_llr() / Constellation.llr()", "float32")
    out = llr_any(constellation, t, nv, method)
    return to_host_like(out, placed.kind)
