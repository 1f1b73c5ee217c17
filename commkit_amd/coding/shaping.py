"""Distribution matching: :class:`CCDM` and probabilistic amplitude shaping, :class:`PAS`.

Restates ``commkit/coding/shaping.py`` (commkit 3.11.0).  A constant-
composition distribution matcher maps ``k`` uniform bits to a block of
``num_symbols`` constellation points holding exactly :attr:`CCDM.composition`
of each point: the KL-optimal quantization of the pmf to ``1 /
num_symbols``.  Each block of bits is a binary fraction; the sequences of the
composition partition ``[0, 1)`` by a 64-bit range coder (floor rounding,
the remainder to the most frequent point) and the matcher emits the sequence
whose interval holds the fraction.  ``k`` is small enough that no interval
exceeds ``2^-k`` even with the rounding, so the map is injective and the
dematcher inverts it exactly, flagging sequences that no payload produces.

The range coder runs in the HIP kernels ``ck_ccdm_match`` and
``ck_ccdm_dematch`` (one lane per block, parallel over blocks); the tables
here (composition, ``k``, buffer size, label lookup) are host arithmetic,
computed once per object.

:class:`PAS` puts a CCDM on the amplitudes of one real dimension and a
systematic code's parity (and any information bits left over) on the signs;
:func:`~commkit_amd.coding.encode` and :func:`~commkit_amd.coding.decode`
take both classes like any code.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from functools import cached_property
from typing import Any

import numpy as np

from .. import _native
from .._array import require_int
from ..backend import device_tables
from ..mapping import Constellation
from ..mapping.constellation import pack_bits

__all__ = ["CCDM", "PAS"]

# Range-coder precision (commkit/coding/shaping.py:60-63): the range is
# renormalized to [2^31, 2^32).
_RANGE_BITS = 32
_MAX_BLOCK = 1 << 30
# Points a matcher's constellation may have (the kernels' limit).
_MAX_POINTS = 64


def _composition(pmf: np.ndarray, n: int) -> np.ndarray:
    """KL-optimal counts (sum ``n``) for ``pmf``: ``floor(n p)``, then one
    point at a time where the divergence grows least, the first on ties
    (``commkit/coding/shaping.py:201-214``)."""
    p = np.asarray(pmf, dtype=np.float64)
    c = np.floor(n * p).astype(np.int64)
    for _ in range(n - int(c.sum())):
        with np.errstate(divide="ignore", invalid="ignore"):
            grow = (c + 1) * np.log((c + 1) / (n * p)) - np.where(
                c > 0, c * np.log(c / (n * p)), 0.0
            )
        grow[p == 0] = np.inf
        c[int(np.argmin(grow))] += 1
    return c


@dataclass(frozen=True, eq=False)
class CCDM:
    """Constant-composition distribution matcher onto a constellation
    (``commkit/coding/shaping.py:69-198``).

    Parameters
    ----------
    constellation : Constellation
        The target: its pmf (uniform if none), points and labels; at most 64
        points.
    num_symbols : int
        Points per block, ``1 <= num_symbols <= 2^30``.

    ``n = num_symbols * bits_per_symbol`` transmitted (label) bits carry
    ``k`` information bits per block.
    """

    constellation: Constellation
    num_symbols: int

    def __post_init__(self) -> None:
        if not isinstance(self.constellation, Constellation):
            raise TypeError(
                "CCDM(): constellation must be a Constellation, got "
                f"{type(self.constellation).__name__}."
            )
        n = require_int(self.num_symbols, "num_symbols", low=1, high=_MAX_BLOCK)
        object.__setattr__(self, "num_symbols", n)
        if self.constellation.order > _MAX_POINTS:
            raise ValueError(
                f"CCDM(): {self.constellation.order} points; the matcher takes at most "
                f"{_MAX_POINTS}."
            )
        if self.k < 1:
            raise ValueError(
                f"CCDM(): a block of {n} points carries no information bit; use a "
                "longer block."
            )

    # -- value semantics ---------------------------------------------------

    def __eq__(self, other: object) -> bool:
        if self is other:
            return True
        if not isinstance(other, CCDM):
            return NotImplemented
        return self.num_symbols == other.num_symbols and self.constellation == other.constellation

    def __hash__(self) -> int:
        return hash((self.constellation, self.num_symbols))

    def __repr__(self) -> str:
        return f"CCDM({self.constellation!r}, num_symbols={self.num_symbols})"

    # -- sizes -------------------------------------------------------------

    @property
    def n(self) -> int:
        """Transmitted bits per block: ``num_symbols * bits_per_symbol``."""
        return self.num_symbols * self.constellation.bits_per_symbol

    @property
    def k(self) -> int:
        """Information bits per block."""
        return self._k

    @property
    def rate(self) -> float:
        return float(self.k / self.n)

    @cached_property
    def composition(self) -> np.ndarray:
        """Points of each kind per block, ``(M,)`` int64 summing to ``num_symbols``."""
        c = _composition(self._pmf, self.num_symbols)
        c.setflags(write=False)
        return c

    @cached_property
    def _pmf(self) -> np.ndarray:
        pmf = self.constellation.pmf
        order = self.constellation.order
        return np.full(order, 1.0 / order) if pmf is None else np.asarray(pmf)

    @cached_property
    def _log2_sequences(self) -> float:
        """``log2`` of the number of sequences of the composition."""
        lg = math.lgamma(self.num_symbols + 1) - sum(math.lgamma(int(c) + 1)
                                                     for c in self.composition)  # fmt: skip
        return lg / math.log(2)

    @cached_property
    def _k(self) -> int:
        """``floor(log2 #sequences - slack - margin)``: the floor rounding lets the
        most frequent point's share grow by at most ``(a - 1) / range`` per
        position (``a`` points in use), which costs ``slack`` bits over the
        block; ``margin`` covers lgamma's rounding
        (``commkit/coding/shaping.py:150-160``)."""
        log2_m = self._log2_sequences
        a = int(np.count_nonzero(self.composition))
        slack = self.num_symbols * math.log2(1.0 + a * (a - 1) / 2.0 ** (_RANGE_BITS - 1))
        margin = 1e-9 * max(1.0, log2_m)
        return max(0, math.floor(log2_m - slack - margin))

    @cached_property
    def _buffer_bits(self) -> int:
        """Fractional bits of the dematcher's interval start, with margin
        (``commkit/coding/shaping.py:162-167``)."""
        n = self.num_symbols
        shrink = -n * math.log2(1.0 - n / 2.0 ** (_RANGE_BITS - 1))
        return int(math.ceil(self._log2_sequences + shrink)) + 2 * _RANGE_BITS + 8

    @cached_property
    def _label_index(self) -> np.ndarray:
        """Point index of each packed label, ``(M,)`` int64."""
        lut = np.empty(self.constellation.order, dtype=np.int64)
        lut[pack_bits(self.constellation.bit_labels)] = np.arange(self.constellation.order)
        return lut

    # -- the range coder (device) -------------------------------------------

    def _match(self, bits: Any) -> Any:
        """``(R, k)`` int8 device bits to ``(R, num_symbols)`` int32 point indices."""
        import torch

        lib = _native.lib()
        bits = bits.to(torch.int8).contiguous()
        rows = int(bits.shape[0])
        out = torch.empty((rows, self.num_symbols), dtype=torch.int32, device=bits.device)
        comp = np.ascontiguousarray(self.composition, dtype=np.int64)
        with torch.cuda.device(bits.device):
            _native.check(lib.ck_ccdm_match(
                bits.data_ptr() if rows else None, rows, self.k, comp.ctypes.data, comp.size,
                self.num_symbols, out.data_ptr() if rows else None,
                _native.stream_of(bits.device)))  # fmt: skip
        return out

    def _dematch(self, symbols: Any) -> tuple[Any, Any]:
        """``(R, num_symbols)`` device point indices to ``(R, k)`` int8 bits and
        ``(R,)`` bool validity."""
        import torch

        lib = _native.lib()
        sym = symbols.to(torch.int32).contiguous()
        rows = int(sym.shape[0])
        dev = sym.device
        bits = torch.empty((rows, self.k), dtype=torch.int8, device=dev)
        valid = torch.empty(rows, dtype=torch.uint8, device=dev)
        if rows == 0:
            return bits, valid.bool()
        comp = np.ascontiguousarray(self.composition, dtype=np.int64)
        ws_bytes = lib.ck_ccdm_dematch_workspace_bytes(rows, self._buffer_bits)
        if ws_bytes < 0:
            raise RuntimeError("ck_ccdm_dematch_workspace_bytes failed")
        ws = torch.empty(max(int(ws_bytes), 1), dtype=torch.uint8, device=dev)
        with torch.cuda.device(dev):
            _native.check(lib.ck_ccdm_dematch(
                sym.data_ptr(), rows, self.num_symbols, comp.ctypes.data, comp.size, self.k,
                self._buffer_bits, bits.data_ptr(), valid.data_ptr(), ws.data_ptr(),
                _native.stream_of(dev)))  # fmt: skip
        return bits, valid.bool()

    # -- labels ---------------------------------------------------------------

    def _encode(self, bits: Any) -> Any:
        """``(R, k)`` device bits to ``(R, n)`` int8 label bits."""
        import torch

        symbols = self._match(bits)
        (labels,) = device_tables(self, "ccdm_labels", bits.device,
                                  lambda: (self.constellation.bit_labels.astype(np.int8),))  # fmt: skip
        return labels[symbols.to(torch.int64)].reshape(bits.shape[0], self.n)

    def _dematch_labels(self, labels: Any) -> tuple[Any, Any]:
        """Hard ``(R, n)`` device label bits to ``(R, k)`` bits and validity."""
        import torch

        m = self.constellation.bits_per_symbol
        rows = labels.shape[0]
        weights = (1 << torch.arange(m - 1, -1, -1, device=labels.device, dtype=torch.int64))
        packed = (labels.reshape(rows, self.num_symbols, m).to(torch.int64) * weights).sum(-1)
        (lut,) = device_tables(self, "ccdm_label_index", labels.device,
                               lambda: (self._label_index,))  # fmt: skip
        return self._dematch(lut[packed])


@dataclass(frozen=True, eq=False)
class PAS:
    """Probabilistic amplitude shaping: a CCDM on the amplitudes, a systematic
    code for the signs (``commkit/coding/shaping.py:249-389``).

    Parameters
    ----------
    code : LDPCCode
        A systematic code of rate at least ``(m - 1) / m`` (``m`` bits per real
        dimension) with ``n`` a multiple of ``2 m`` (QAM) or ``m`` (PAM).
    constellation : Constellation
        Gray square QAM (I bits first) or bipolar PAM with a symmetric pmf.

    Per codeword of ``n_c`` bits there are ``N = n_c / m`` real dimensions: a
    CCDM of ``N`` amplitudes maps ``k_dm`` bits to ``(m - 1) N`` amplitude
    bits, the code encodes them with ``k_c - (m - 1) N`` uniform bits, and
    those uniform bits and the parity are the signs; ``k = k_dm + k_c - (m -
    1) N``.  Transmitted bit ``j`` is codeword bit ``permutation[j]``:
    per dimension the sign, then its amplitude bits.
    """

    code: Any
    constellation: Constellation

    def __post_init__(self) -> None:
        from .ldpc import LDPCCode

        if not isinstance(self.code, LDPCCode):
            raise TypeError(
                f"PAS(): code must be a code object (LDPCCode), got {type(self.code).__name__}."
            )
        if not isinstance(self.constellation, Constellation):
            raise TypeError("PAS(): constellation must be a Constellation.")
        m = self._layout[0]
        dims = 2 if self.constellation.is_complex else 1
        if self.code.n % (dims * m):
            raise ValueError(
                f"PAS(): the code length {self.code.n} must be a multiple of "
                f"{dims * m} ({dims} dimension(s) of {m} bits)."
            )
        if self._sign_info < 0:
            raise ValueError(
                f"PAS(): the code rate {self.code.rate:.4f} is below (m - 1) / m = "
                f"{(m - 1) / m:.4f}: its parity cannot all be signs."
            )
        _ = self.matcher

    def __eq__(self, other: object) -> bool:
        if self is other:
            return True
        if not isinstance(other, PAS):
            return NotImplemented
        return self.code == other.code and self.constellation == other.constellation

    def __hash__(self) -> int:
        return hash((self.code, self.constellation))

    @property
    def n(self) -> int:
        """Transmitted bits per codeword."""
        return int(self.code.n)

    @property
    def k(self) -> int:
        """Information bits per codeword."""
        return int(self.matcher.k + self._sign_info)

    @property
    def rate(self) -> float:
        return float(self.k / self.n)

    def __repr__(self) -> str:
        return f"PAS({self.code!r}, {self.constellation!r})"

    @cached_property
    def _layout(self) -> tuple[int, Constellation]:
        """``(m, amplitudes)``: bits per real dimension and the amplitude
        alphabet (the positive levels, their labels without the sign bit and
        twice their prior) (``commkit/coding/shaping.py:318-365``)."""
        c = self.constellation
        if c.is_complex:
            axes = c._axes
            m = c.bits_per_symbol // 2
            if (
                axes is None
                or c.bits_per_symbol % 2
                or not np.array_equal(axes[0].bits, np.arange(m))
                or not np.array_equal(axes[1].bits, np.arange(m, 2 * m))
                or not np.allclose(axes[0].levels, axes[1].levels)
                or not np.allclose(
                    np.zeros(1) if axes[0].pmf is None else axes[0].pmf,
                    np.zeros(1) if axes[1].pmf is None else axes[1].pmf,
                )
            ):
                raise ValueError(
                    "PAS(): needs square QAM with the I bits first and the same "
                    "levels and prior on both axes, as Constellation.qam builds."
                )
            levels, labels, pmf = axes[0].levels, axes[0].labels, axes[0].pmf
        else:
            order = np.argsort(c.points)
            levels, labels = c.points[order], c.bit_labels[order]
            pmf = None if c.pmf is None else c.pmf[order]
            m = c.bits_per_symbol
        size = levels.size
        mirror = labels[::-1]
        if (
            m < 2
            or not np.allclose(levels, -levels[::-1])
            or not np.array_equal(labels[:, 0], (levels > 0).astype(labels.dtype))
            or not np.array_equal(labels[:, 1:], mirror[:, 1:])
            or (pmf is not None and not np.allclose(pmf, pmf[::-1]))
        ):
            raise ValueError(
                "PAS(): needs at least two bits per real dimension, a sign bit "
                "first, the same amplitude label for +a and -a and a symmetric "
                "prior (Gray QAM or bipolar PAM; unipolar PAM has no sign)."
            )
        positive = slice(size // 2, size)
        amp_pmf = None if pmf is None else 2 * np.asarray(pmf)[positive]
        amplitudes = Constellation(np.asarray(levels)[positive],
                                   bit_labels=np.asarray(labels)[positive, 1:], pmf=amp_pmf)  # fmt: skip
        return m, amplitudes

    @cached_property
    def matcher(self) -> CCDM:
        """The CCDM of the amplitudes of one codeword."""
        m, amplitudes = self._layout
        return CCDM(amplitudes, num_symbols=self.code.n // m)

    @cached_property
    def _sign_info(self) -> int:
        """Uniform information bits that become signs."""
        m = self._layout[0]
        return int(self.code.k - (m - 1) * (self.code.n // m))

    @cached_property
    def permutation(self) -> np.ndarray:
        """``(n,)`` int64: transmitted bit ``j`` is codeword bit ``permutation[j]``
        (``commkit/coding/shaping.py:379-389``)."""
        m = self._layout[0]
        dims = self.code.n // m
        info = np.asarray(self.code._encoder.info, dtype=np.int64)
        parity = np.setdiff1d(np.arange(self.code.n), info)
        amplitude = info[: (m - 1) * dims].reshape(dims, m - 1)
        sign = np.concatenate([info[(m - 1) * dims:], parity])
        perm = np.concatenate([sign[:, None], amplitude], axis=1).ravel().astype(np.int64)
        perm.setflags(write=False)
        return perm

    def _tables(self, device: Any) -> tuple[Any, Any]:
        """The permutation and its inverse on ``device``."""
        return device_tables(self, "pas_perm", device,
                             lambda: (self.permutation, np.argsort(self.permutation)))  # fmt: skip

    def _encode(self, bits: Any) -> Any:
        """``(R, k)`` device bits to ``(R, n)`` int8 transmitted bits."""
        import torch

        from .ldpc import _encode as ldpc_encode

        ccdm = self.matcher
        amplitude_bits = ccdm._encode(bits[:, : ccdm.k])
        info = torch.cat([amplitude_bits, bits[:, ccdm.k:].to(torch.int8)], dim=1)
        words = ldpc_encode(self.code, info)
        perm, _ = self._tables(bits.device)
        return words[:, perm]

    def _decode(self, llrs: Any, method: Any) -> tuple[Any, Any, Any, Any]:
        """``(R, n)`` float32 device LLRs of the transmitted bits (with the
        prior) to ``(bits, app, valid, iterations)``: the code's decoder on the
        de-permuted LLRs, then the dematcher on its amplitude bits."""
        import torch

        from .ldpc import _decode as ldpc_decode

        ccdm = self.matcher
        perm, inv = self._tables(llrs.device)
        bits, app, valid, iters = ldpc_decode(self.code, llrs[:, inv], method)
        dm_bits, dm_valid = ccdm._dematch_labels(bits[:, : ccdm.n])
        out = torch.cat([dm_bits, bits[:, ccdm.n:]], dim=1)
        return out, app[:, perm], valid & dm_valid, iters
