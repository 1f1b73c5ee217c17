"""Scoring of the receive tail: :func:`ber`, :func:`fer`, :func:`gmi`, :func:`ngmi`.

Signatures and conventions of ``commkit/metrics/error_rates.py:30-181`` and
``commkit/metrics/information.py:34-247``: bits (0/1) or LLRs ``(N,)`` or
``(C, N)`` against the transmitted bits of the same shape; a float for 1-D
input, a float64 ``(C,)`` array for 2-D input.  A
:class:`~commkit_amd.coding.DecodeResult` is scored against the reference's
information bits (a Signal, its :class:`~commkit_amd.core.Reference`, or the
bits themselves), in blocks of the code's ``k`` bits
(``commkit/metrics/_common.py:65-87``).  The counts and the GMI loss are
reductions in the HIP kernels ``ck_bit_errors`` and ``ck_gmi_loss`` on the
data's device; only the per-row results come back to the host.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from . import _native
from ._array import require_int
from .backend import dispatch
from .coding.base import DecodeResult
from .core import Reference, Signal

__all__ = ["awgn_gmi", "ber", "fer", "gmi", "ngmi"]


def _report(values: np.ndarray, one_d: bool) -> float | np.ndarray:
    return float(values[0]) if one_d else values


def _decoded(bits: Any, reference: Any, name: str) -> tuple[Any, Any, Any] | None:
    """(decoded bits, information bits, code) for a DecodeResult, else None
    (``commkit/metrics/_common.py:65-87``)."""
    if not isinstance(bits, DecodeResult):
        return None
    ref = reference.reference if isinstance(reference, Signal) else reference
    if isinstance(ref, Reference):
        info = ref.info_bits if ref.info_bits is not None else ref.bits
        if info is None:
            raise ValueError(f"{name}: the reference has no bits.")
        return bits.bits, info, ref.code
    if ref is None:
        raise ValueError(f"{name}: a DecodeResult needs the transmitted Signal, its Reference "
                         "or the information bits as reference.")  # fmt: skip
    return bits.bits, ref, None


def _signal_reference(signal: Signal | None, reference: Any, field: str, name: str) -> Any:
    """The explicit reference, else the Signal's ``reference.<field>``."""
    if reference is not None:
        return reference
    ref = None if signal is None else signal.reference
    value = None if ref is None else getattr(ref, field)
    if value is None:
        raise ValueError(f"{name} needs reference {field}: pass reference= or a Signal whose "
                         f"reference has {field}.")  # fmt: skip
    return value


def _rows(x: Any, name: str) -> tuple[Any, bool]:
    if x.ndim not in (1, 2):
        raise ValueError(f"{name}: expected (N,) or (C, N), got {x.ndim}-D.")
    return (x.reshape(1, -1) if x.ndim == 1 else x), x.ndim == 1


def _paired_bits(rx: Any, reference: Any, skip: int, name: str) -> tuple[Any, Any, bool]:
    """Both as (rows, N - skip) int8 0/1 tensors on the received data's device."""
    import torch

    tx = dispatch(reference, "reference").tensor.to(rx.device)
    if tuple(rx.shape) != tuple(tx.shape):
        raise ValueError(f"{name}: shape mismatch: received {tuple(rx.shape)} != reference "
                         f"{tuple(tx.shape)}.")  # fmt: skip
    rx2, one_d = _rows(rx, name)
    tx2, _ = _rows(tx, name)
    if skip:
        rx2, tx2 = _skip(rx2, skip, name), _skip(tx2, skip, name)
    return ((rx2 != 0).to(torch.int8).contiguous(), (tx2 != 0).to(torch.int8).contiguous(), one_d)


def _skip(x: Any, n: int, name: str) -> Any:
    x = x[..., n:]
    if x.shape[-1] == 0:
        raise ValueError(f"{name}: nothing to measure (no symbols after skipping).")
    return x


def _errors(rx: Any, tx: Any, block: int | None) -> tuple[np.ndarray, np.ndarray | None]:
    """Bit errors and blocks of ``block`` bits in error per row, from
    ``ck_bit_errors`` (``block`` None: bit errors only)."""
    import torch

    rows, n = int(rx.shape[0]), int(rx.shape[1])
    bit_err = torch.empty(rows, dtype=torch.int64, device=rx.device)
    block_err = None if block is None else torch.empty(rows, dtype=torch.int64, device=rx.device)
    with torch.cuda.device(rx.device):
        _native.check(_native.lib().ck_bit_errors(
            rx.data_ptr(), tx.data_ptr(), rows, n, n if block is None else block,
            bit_err.data_ptr(), None if block_err is None else block_err.data_ptr(),
            _native.stream_of(rx.device)))  # fmt: skip
    return bit_err.cpu().numpy(), None if block_err is None else block_err.cpu().numpy()


def _hard_bits(signal: Signal, constellation: Any, name: str) -> tuple[Any, Any]:
    """Minimum-distance decisions of a Signal's symbols, as their labels.

    The sign of the max-log LLR of the constellation without its prior is the
    label bit of the nearest point (max-log keeps, for each bit value, only
    the nearest point carrying it; the overall nearest point is one of the
    two)."""
    import torch

    from .core.signal import signal_input
    from .mapping import Constellation
    from .mapping.demap import llr_any

    x, c = signal_input(signal, constellation, name)
    plain = c if c.pmf is None else Constellation(c.points, bit_labels=c.bit_labels)
    t = dispatch(x, "symbols").tensor
    return (llr_any(plain, t, 1.0, "maxlog") < 0).to(torch.int8), c


def ber(bits: Any, reference: Any = None, *, constellation: Any = None,
        num_skip_symbols: int = 0) -> float | np.ndarray:
    """Bit error rate.

    ``bits``: received bits (0/1), ``(N,)`` or ``(C, N)``; a Signal of
    received symbols at one sample per symbol, decided against the
    constellation (minimum distance) and compared with ``reference.bits``
    (pre-FEC BER); or a DecodeResult, compared with the reference's
    information bits (post-FEC BER).  ``num_skip_symbols`` leading symbols
    (``bits_per_symbol`` bits each) are left out.  Raises ValueError on a
    missing reference or constellation, a shape mismatch or no bits
    (``commkit/metrics/error_rates.py:30-105``).
    """
    name = "ber()"
    post = _decoded(bits, reference, name)
    signal = None
    if post is not None:
        if num_skip_symbols:
            raise ValueError(f"{name}: num_skip_symbols does not apply to decoded bits.")
        bits, reference, _ = post
    skip = require_int(num_skip_symbols, f"{name}: num_skip_symbols", low=0)
    if isinstance(bits, Signal):
        signal = bits
        rx, c = _hard_bits(bits, constellation, name)
        reference = _signal_reference(signal, reference, "bits", name)
    else:
        if reference is None:
            raise ValueError(f"{name} needs reference bits: pass reference=.")
        c = constellation
        if skip and c is None:
            raise ValueError(f"{name}: num_skip_symbols needs the constellation.")
        rx = dispatch(bits, "bits").tensor
    if skip:
        skip *= c.bits_per_symbol
    rx2, tx2, one_d = _paired_bits(rx, reference, skip, name)
    n = int(rx2.shape[1])
    if n == 0:
        raise ValueError(f"{name}: nothing to measure (no bits).")
    errors, _ = _errors(rx2, tx2, None)
    return _report(errors / n, one_d)


def fer(bits: Any, reference: Any, *, info_bits: int | None = None) -> float | np.ndarray:
    """Block error rate: the fraction of blocks of ``info_bits`` bits with any
    bit in error.  With a DecodeResult and a Signal or Reference holding the
    code, ``info_bits`` defaults to the code's ``k`` and must agree with it
    (``commkit/metrics/error_rates.py:108-181``)."""
    name = "fer()"
    post = _decoded(bits, reference, name)
    if post is not None:
        bits, reference, code = post
        k = None if code is None else int(code.k)
        if info_bits is None:
            info_bits = k
        elif k is not None and info_bits != k:
            raise ValueError(f"{name}: info_bits={info_bits} disagrees with the reference's "
                             f"code (k = {k}).")  # fmt: skip
    if info_bits is None:
        raise ValueError(f"{name}: give info_bits, or a DecodeResult and the Signal or "
                         "Reference that holds the code.")  # fmt: skip
    info_bits = require_int(info_bits, f"{name}: info_bits", low=1)
    rx2, tx2, one_d = _paired_bits(dispatch(bits, "bits").tensor, reference, 0, name)
    total = int(rx2.shape[1])
    if total == 0:
        raise ValueError(f"{name}: nothing to measure (no bits).")
    if total % info_bits:
        raise ValueError(f"{name}: the last axis ({total}) is not a multiple of info_bits "
                         f"({info_bits}).")  # fmt: skip
    _, blocks = _errors(rx2, tx2, info_bits)
    return _report(blocks / (total // info_bits), one_d)


def _data_aided_gain(rx: Any, tx: Any) -> Any:
    """``|<r s*>| / <|s|^2>`` per row, accumulated in float64, 1 where there
    is no reference power or correlation (``commkit/math.py:235-258``)."""
    import torch

    wide = torch.complex128 if rx.is_complex() or tx.is_complex() else torch.float64
    r, s = rx.to(wide), tx.to(wide)
    num = (r * s.conj()).mean(dim=-1, keepdim=True).abs()
    den = (s.abs() ** 2).mean(dim=-1, keepdim=True)
    g = num / den
    ok = (den > 0) & (g > 0)
    return torch.where(ok, g, torch.ones_like(g)).to(rx.real.dtype)


def _gmi_loss(llrs: Any, reference: Any, constellation: Any, noise_var: Any, esn0_db: Any,
              method: str, num_skip_symbols: int, name: str) -> tuple[np.ndarray, Any, bool]:
    """``E[log2(1 + e^{-(1 - 2c) L})]`` per row (``commkit/metrics/
    information.py:34-89``), from ``ck_gmi_loss``; and the constellation."""
    import torch

    from .core.signal import signal_input
    from .mapping.demap import check_method, llr_any, noise_variance

    skip = require_int(num_skip_symbols, f"{name}: num_skip_symbols", low=0)
    signal = llrs if isinstance(llrs, Signal) else None
    if signal is not None:
        x, c = signal_input(signal, constellation, name)
        if noise_var is None and esn0_db is None:
            raise ValueError(f"{name} needs noise_var or esn0_db to compute LLRs of a Signal.")
        check_method(method)
        t = dispatch(x, "symbols").tensor
        ref_sym = None if signal.reference is None else signal.reference.symbols
        if ref_sym is not None:
            s = dispatch(ref_sym, "reference symbols").tensor.to(t.device)
            if tuple(s.shape) != tuple(t.shape):
                raise ValueError(f"{name}: shape mismatch: received {tuple(t.shape)} != "
                                 f"reference {tuple(s.shape)}.")  # fmt: skip
            g = _data_aided_gain(_skip(t, skip, name), _skip(s, skip, name))
            t = t / (g if t.ndim > 1 else g.reshape(()))
        nv = noise_variance(noise_var, esn0_db, c, not t.is_complex(), name)
        x = llr_any(c, t, nv, method)
    else:
        c = constellation
        if c is None:
            raise ValueError(f"{name}: constellation is required for LLR arrays.")
        if noise_var is not None or esn0_db is not None:
            raise ValueError(f"{name}: noise_var and esn0_db are for Signal input; LLR arrays "
                             "already contain the noise.")  # fmt: skip
        x = dispatch(llrs, "llrs").tensor
    reference = _signal_reference(signal, reference, "bits", name)
    tx = dispatch(reference, "reference").tensor.to(x.device)
    if tuple(tx.shape) != tuple(x.shape):
        raise ValueError(f"{name}: shape mismatch: received {tuple(x.shape)} != reference "
                         f"{tuple(tx.shape)}.")  # fmt: skip
    k = c.bits_per_symbol
    x2, one_d = _rows(x, name)
    tx2, _ = _rows(tx, name)
    if int(x2.shape[1]) % k:
        raise ValueError(f"{name}: the last axis ({int(x2.shape[1])}) is not a multiple of "
                         f"bits_per_symbol ({k}).")  # fmt: skip
    if skip:
        x2, tx2 = _skip(x2, skip * k, name), _skip(tx2, skip * k, name)
    if int(x2.shape[1]) == 0:
        raise ValueError(f"{name}: nothing to measure (no LLRs).")
    x2 = x2.to(torch.float32).contiguous()
    tx2 = (tx2 != 0).to(torch.int8).contiguous()
    rows, n = int(x2.shape[0]), int(x2.shape[1])
    loss = torch.empty(rows, dtype=torch.float64, device=x2.device)
    with torch.cuda.device(x2.device):
        _native.check(_native.lib().ck_gmi_loss(x2.data_ptr(), tx2.data_ptr(), rows, n,
                                                loss.data_ptr(), _native.stream_of(x2.device)))  # fmt: skip
    return loss.cpu().numpy(), c, one_d


def gmi(llrs: Any, reference: Any = None, *, constellation: Any = None, noise_var: Any = None,
        esn0_db: Any = None, method: str = "maxlog",
        num_skip_symbols: int = 0) -> float | np.ndarray:
    """Generalized mutual information in bits per symbol,
    ``H(X) - sum_b E[log2(1 + e^{-(1 - 2 c_b) L_b})]`` with the entropy of
    the constellation's prior (the LLRs must include the prior, as
    compute_llr's do).  A Signal's symbols are divided by their data-aided
    gain against ``reference.symbols`` (when present) and demapped with
    ``noise_var`` or ``esn0_db`` and ``method``
    (``commkit/metrics/information.py:92-181``)."""
    loss, c, one_d = _gmi_loss(llrs, reference, constellation, noise_var, esn0_db, method,
                               num_skip_symbols, "gmi()")  # fmt: skip
    k = c.bits_per_symbol
    return _report(k * (c.entropy / k - loss), one_d)


def ngmi(llrs: Any, reference: Any = None, *, constellation: Any = None, noise_var: Any = None,
         esn0_db: Any = None, method: str = "maxlog",
         num_skip_symbols: int = 0) -> float | np.ndarray:
    """Normalized GMI, ``1 - E[log2(1 + e^{-(1 - 2c) L})]`` averaged over the
    bits: 1 for error-free LLRs, the same for uniform and shaped
    constellations at equal LLR quality (``commkit/metrics/information.py:
    184-247``)."""
    loss, _, one_d = _gmi_loss(llrs, reference, constellation, noise_var, esn0_db, method,
                               num_skip_symbols, "ngmi()")  # fmt: skip
    return _report(1.0 - loss, one_d)



#: Gauss-Hermite nodes per noise axis of :func:`awgn_gmi` (the reference's
#: order; doubling it moves no rate by more than a few 1e-7 bit).
AWGN_NODES = 128


def awgn_gmi(constellation: Any, *, esn0_db: Any) -> float | np.ndarray:
    """GMI (bit-metric decoding rate) of ``constellation`` over AWGN, in bits
    per symbol: ``H(X) - sum_b H(B_b | Y)`` with exact bit LLRs including the
    prior (``commkit/metrics/awgn.py:477-523``).

    The expectation over the noise is Gauss-Hermite quadrature of order
    :data:`AWGN_NODES` per axis (a product rule in the plane for complex
    points): every point plus every noise node is demapped by the library's
    exact demapper kernel on the GPU, and the loss terms are summed with the
    point prior and the node weights in float64.  ``Es`` is the average
    energy under the prior; complex points see complex noise of variance
    ``N0``, real points real noise of variance ``N0 / 2``.  A float for a
    scalar ``esn0_db``, an array of its shape otherwise.  Raises TypeError for
    a non-Constellation, ValueError for a non-finite Es/N0.
    """
    import torch

    from .mapping import Constellation
    from .mapping.demap import llr_any, noise_variance

    name = "awgn_gmi()"
    if not isinstance(constellation, Constellation):
        raise TypeError(f"{name}: constellation must be a Constellation, got "
                        f"{type(constellation).__name__}.")  # fmt: skip
    es = np.asarray(esn0_db, dtype=np.float64)
    if not np.all(np.isfinite(es)):
        raise ValueError(f"{name}: esn0_db must be finite.")
    c = constellation
    if not torch.cuda.is_available():
        raise RuntimeError(f"{name}: the demapper runs on a GPU; none is visible.")
    dev = torch.device("cuda", torch.cuda.current_device())
    k = c.bits_per_symbol
    pmf = np.full(c.order, 1.0 / c.order) if c.pmf is None else np.asarray(c.pmf)
    t, w = np.polynomial.hermite.hermgauss(AWGN_NODES)
    if c.is_complex:
        nodes = (t[:, None] + 1j * t[None, :]).reshape(-1)
        weights = (w[:, None] * w[None, :]).reshape(-1) / np.pi
    else:
        nodes, weights = t.astype(np.complex128), w / np.sqrt(np.pi)
    pts = torch.from_numpy(np.array(c.points, dtype=np.complex128)).to(dev)
    node_t = torch.from_numpy(nodes).to(dev)
    # weight of each (point, node) pair, and the sign of each label bit
    pair_w = torch.from_numpy(np.outer(pmf, weights)).to(dev)
    sign = torch.from_numpy(1.0 - 2.0 * np.asarray(c.bit_labels, np.float64)).to(dev)
    out = np.empty(es.size)
    for j, e in enumerate(es.reshape(-1)):
        nv = noise_variance(None, float(e), c, not c.is_complex, name)
        # e^{-t^2} quadrature: noise of variance s2 per axis is sqrt(2 s2) t
        scale = float(np.sqrt(2.0 * (nv / 2.0 if c.is_complex else nv)))
        y = pts[:, None] + scale * node_t[None, :]
        y = y.to(torch.complex64) if c.is_complex else y.real.to(torch.float32)
        llrs = llr_any(c, y.reshape(-1), nv, "exact").reshape(c.order, -1, k)
        x = -llrs.to(torch.float64) * sign[:, None, :]
        loss = torch.logaddexp(torch.zeros((), dtype=torch.float64, device=dev), x).sum(dim=2)
        out[j] = c.entropy - float((pair_w * loss).sum()) / np.log(2.0)
    return float(out[0]) if es.ndim == 0 else out.reshape(es.shape)
