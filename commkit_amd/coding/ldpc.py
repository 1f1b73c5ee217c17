"""LDPC codes: :class:`LDPCCode`, :class:`MinSum` and :class:`SumProduct`.

A code is its sparse parity-check matrix ``H`` (``m x N``), optionally with
the last ``punctured`` bits of each codeword not transmitted, as
``commkit/coding/ldpc.py:53-125``.  Built from ``H``
(:meth:`LDPCCode.from_parity_check`) or from a quasi-cyclic base matrix
(:meth:`LDPCCode.quasi_cyclic`, ``ldpc.py:182-214``: shift ``s`` puts row
``r``'s one at column ``(r + s) mod Z``); a quasi-cyclic code keeps its base
matrix, so the decoder computes columns from shifts.

Decoding schedule (``ldpc.py:283-296,422-442``): rows are coloured greedily
in index order, each taking the first layer none of whose rows shares a
bit, and stably sorted by colour.  For a quasi-cyclic code every row of a
block row takes the same colour (each nonzero block is a permutation, so
earlier block rows sharing a block column use every one of its bits), so a
layer is a union of whole block rows and the block rows run in colour
order.

Encoding is systematic, information bits first (``ldpc.py:445-536``):
a dual-diagonal parity part is accumulated, an invertible one (``m <=
8192``) is inverted over GF(2), otherwise a dense systematic generator comes
from row reduction with the rightmost pivots.

Decoding runs ``ck_ldpc_decode`` of ``libck_hip.so``: layered (the
reference's schedule) or flooding (an addition: every check uses the
previous iteration's APP, then ``APP = L + sum R`` in ascending check
order), normalized/offset min-sum or sum-product, syndrome early stop.
"""

from __future__ import annotations

import ctypes
import logging
from dataclasses import dataclass
from functools import cached_property
from typing import Any, NamedTuple

import numpy as np

from .. import _native
from .._array import require_finite, require_int
from ..backend import device_tables
from .gf2 import inverse, row_reduce

__all__ = ["LDPCCode", "MinSum", "SumProduct"]

logger = logging.getLogger("commkit_amd")

# Largest parity part inverted densely (m x m bits).
_MAX_INVERSE = 8192
# Largest m * N elimination of the dense-generator fallback.
_MAX_DENSE = 1 << 27
# Device workspace per decode launch (bytes); larger batches are chunked.
_GPU_WORKSPACE = 1 << 31

_SCHEDULES = ("layered", "flooding")


@dataclass(frozen=True, eq=False)
class LDPCCode:
    """LDPC code given by the edges of its sparse parity-check matrix.

    Parameters
    ----------
    checks, variables : array_like
        Row and column of each nonzero entry of ``H``, ``(E,)`` integers,
        without duplicates.  Every row has at least two entries.
    num_checks, num_variables : int
        ``m`` and ``N`` (the mother codeword length).
    punctured : int, default 0
        Trailing codeword bits not transmitted (parity bits); ``n = N - punctured``.
    """

    checks: np.ndarray
    variables: np.ndarray
    num_checks: int
    num_variables: int
    punctured: int = 0

    def __post_init__(self) -> None:
        rows = np.asarray(self.checks)
        cols = np.asarray(self.variables)
        for name in ("num_checks", "num_variables", "punctured"):
            require_int(getattr(self, name), name)
        m, n = int(self.num_checks), int(self.num_variables)
        if rows.ndim != 1 or rows.shape != cols.shape or rows.size == 0:
            raise ValueError("checks and variables must be equal-length 1-D arrays.")
        if rows.dtype.kind not in "iu" or cols.dtype.kind not in "iu":
            raise TypeError(
                f"checks and variables must be integers, got {rows.dtype} and {cols.dtype}."
            )
        if m < 1 or n <= m:
            raise ValueError(f"need 1 <= num_checks < num_variables, got {m} and {n}.")
        if rows.min() < 0 or rows.max() >= m or cols.min() < 0 or cols.max() >= n:
            raise ValueError("checks or variables out of range.")
        order = np.lexsort((cols, rows))
        rows, cols = rows[order].astype(np.int32), cols[order].astype(np.int32)
        flat = rows.astype(np.int64) * n + cols
        if np.any(flat[1:] == flat[:-1]):
            raise ValueError("H has duplicate entries.")
        if np.bincount(rows, minlength=m).min() < 2:
            raise ValueError("every parity check needs at least two bits.")
        p = require_int(self.punctured, "punctured", low=0)
        rows.setflags(write=False)
        cols.setflags(write=False)
        object.__setattr__(self, "checks", rows)
        object.__setattr__(self, "variables", cols)
        object.__setattr__(self, "num_checks", m)
        object.__setattr__(self, "num_variables", n)
        object.__setattr__(self, "punctured", p)
        if self._encoder.info.max(initial=-1) >= n - p:
            raise ValueError("punctured bits must be parity bits.")

    # -- sizes -------------------------------------------------------------------

    @property
    def n(self) -> int:
        """Transmitted bits per codeword."""
        return self.num_variables - self.punctured

    @property
    def k(self) -> int:
        """Information bits per codeword."""
        return int(self._encoder.info.size)

    @property
    def rate(self) -> float:
        return float(self.k / self.n)

    def __repr__(self) -> str:
        return f"LDPCCode(n={self.n}, k={self.k})"

    def __eq__(self, other: object) -> bool:
        if self is other:
            return True
        if not isinstance(other, LDPCCode):
            return NotImplemented
        return (
            self.num_checks == other.num_checks
            and self.num_variables == other.num_variables
            and self.punctured == other.punctured
            and np.array_equal(self.checks, other.checks)
            and np.array_equal(self.variables, other.variables)
        )

    def __hash__(self) -> int:
        h = self.__dict__.get("_hash")
        if h is None:
            h = hash((self.num_checks, self.num_variables, self.punctured,
                      self.checks.tobytes(), self.variables.tobytes()))  # fmt: skip
            self.__dict__["_hash"] = h
        return h

    def parity_check(self) -> Any:
        """``H`` as a ``scipy.sparse.csr_array``, int8, ``(m, N)``."""
        import scipy.sparse

        return scipy.sparse.csr_array(
            (np.ones(self.checks.size, np.int8), (self.checks, self.variables)),
            shape=(self.num_checks, self.num_variables),
        )

    # -- constructors ------------------------------------------------------------

    @classmethod
    def from_parity_check(cls, parity_check: Any, *, punctured: int = 0) -> "LDPCCode":
        """The code of a 0/1 parity-check matrix (dense or scipy.sparse)."""
        import scipy.sparse

        if not scipy.sparse.issparse(parity_check):
            parity_check = np.asarray(parity_check)
        if getattr(parity_check, "ndim", None) != 2:
            raise ValueError(
                f"parity_check must be 2-D, got {getattr(parity_check, 'ndim', None)}-D."
            )
        h = scipy.sparse.coo_array(parity_check)
        h.sum_duplicates()
        h.eliminate_zeros()
        if np.any(h.data != 1):
            raise ValueError("parity_check must contain only 0 and 1.")
        code = cls(h.row, h.col, h.shape[0], h.shape[1], punctured)
        _warn_if_4_cycles(code)
        return code

    @classmethod
    def quasi_cyclic(cls, base: Any, *, lifting: int, punctured: int = 0) -> "LDPCCode":
        """Quasi-cyclic code from a base matrix of circulant shifts (-1: zero block)."""
        b = np.asarray(base)
        z = require_int(lifting, "lifting", low=1)
        if b.ndim != 2 or b.size == 0 or b.dtype.kind not in "iu":
            raise ValueError("base must be a non-empty 2-D integer matrix.")
        if b.min() < -1:
            raise ValueError("base entries must be shifts >= 0 or -1.")
        bi, bj = np.nonzero(b >= 0)
        r = np.arange(z)
        rows = (bi[:, None] * z + r).ravel()
        cols = (bj[:, None] * z + (r + b[bi, bj][:, None]) % z).ravel()
        code = cls(rows, cols, b.shape[0] * z, b.shape[1] * z, punctured)
        base_ro = (np.where(b >= 0, b % z, -1)).astype(np.int64)
        base_ro.setflags(write=False)
        code.__dict__["_qc"] = (base_ro, z)
        _warn_if_4_cycles(code)
        return code

    @classmethod
    def dvbs2(cls, rate: str, *, frame: str = "normal") -> "LDPCCode":
        """DVB-S2 LDPC code (ETSI EN 302 307-1 clause 5.3.2, Annexes B and C),
        without the outer BCH code (``commkit/coding/ldpc.py:217-265``).

        ``frame`` "normal" (n = 64800) or "short" (16200).  Information bit
        ``360 g + j`` checks rows ``(x + j q) mod m`` for each address ``x``
        of group ``g``, ``q = m / 360``; parity bit ``i`` checks rows ``i``
        and ``i + 1`` (the accumulator, so the last parity bit has degree
        1).  Five short labels carry fewer bits than ``n * rate`` (``"1/2"``
        short has k = 7200): read ``code.k``.
        """
        from .standard_tables import DVBS2

        n = {"normal": 64800, "short": 16200}.get(frame)
        if n is None:
            raise ValueError(f"frame must be 'normal' or 'short', got {frame!r}.")
        text = DVBS2.get((frame, rate))
        if text is None:
            rates = [r for f, r in DVBS2 if f == frame]
            raise ValueError(f"rate must be one of {rates} for {frame} frames, got {rate!r}.")
        groups = [np.array(line.split(), dtype=np.int64) for line in text.strip().splitlines()]
        k = 360 * len(groups)
        m = n - k
        q = m // 360
        j = np.arange(360)
        rows = [((x[:, None] + j[None, :] * q) % m).ravel() for x in groups]
        cols = [np.repeat(360 * g + j[None, :], x.size, axis=0).ravel()
                for g, x in enumerate(groups)]  # fmt: skip
        p = np.arange(m)
        rows += [p, p[1:]]
        cols += [k + p, k + p[:-1]]
        return cls(np.concatenate(rows), np.concatenate(cols), m, n)

    @classmethod
    def pon(cls) -> "LDPCCode":
        """LDPC(17280, 14592) of IEEE 802.3ca / ITU-T G.9804.2 Annex B.1: the
        12 x 69 quasi-cyclic mother code with circulants of 256 (N = 17664),
        the last 384 parity bits punctured (``commkit/coding/ldpc.py:267-290``)."""
        from .standard_tables import PON_BASE

        base = np.array([line.split() for line in PON_BASE.strip().splitlines()], dtype=np.int64)
        return cls.quasi_cyclic(base, lifting=256, punctured=384)

    # -- cached structure --------------------------------------------------------

    @cached_property
    def _layer_of_row(self) -> np.ndarray:
        """First-fit layer (colour) of each natural row."""
        m = self.num_checks
        ptr = np.concatenate([[0], np.cumsum(np.bincount(self.checks, minlength=m))])
        return _first_fit(ptr, self.variables, self.num_variables)

    @cached_property
    def _schedule_rows(self) -> np.ndarray:
        """Natural index of each row in schedule order (rows stably sorted by layer)."""
        return np.argsort(self._layer_of_row, kind="stable").astype(np.int64)

    @cached_property
    def _schedule(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """``(row_ptr, cols, layer_ptr)``: CSR of H with rows in layer order."""
        m = self.num_checks
        ptr = np.concatenate([[0], np.cumsum(np.bincount(self.checks, minlength=m))])
        color = self._layer_of_row
        order = self._schedule_rows
        degree = np.diff(ptr)[order]
        row_ptr = np.concatenate([[0], np.cumsum(degree)]).astype(np.int32)
        cols = np.concatenate([self.variables[ptr[r] : ptr[r + 1]] for r in order])
        counts = np.bincount(color)
        layer_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        logger.debug("%r: %d layers.", self, counts.size)
        return row_ptr, cols.astype(np.int32), layer_ptr

    @cached_property
    def _encoder(self) -> "_Encoder":
        return _build_encoder(self)

    @cached_property
    def _qc_view(self) -> dict[str, np.ndarray] | None:
        """The quasi-cyclic tables of ``ck_ldpc_code`` (None for a non-QC code).

        Block rows in schedule order; slots of a block row in ascending block
        column (= ascending column of every row of the block row).
        """
        qc = self.__dict__.get("_qc")
        if qc is None:
            return None
        base, z = qc
        row_ptr, _, _ = self._schedule
        order = self._schedule_rows
        block_order = order[::z] // z  # rows of a block row stay together
        if not np.array_equal(order, (block_order[:, None] * z + np.arange(z)).ravel()):
            return None  # defensive: schedule not block-aligned
        mb, nb = base.shape
        slot_ptr = [0]
        slot_col, slot_shift, slot_pos, slot_row, slot_nat = [], [], [], [], []
        for sbr, br in enumerate(block_order):
            cols = np.flatnonzero(base[br] >= 0)
            slot_col += cols.tolist()
            slot_shift += base[br, cols].tolist()
            slot_pos += list(range(cols.size))
            slot_row += [sbr] * cols.size
            slot_nat += [int(br)] * cols.size
            slot_ptr.append(slot_ptr[-1] + cols.size)
        slot_col_a = np.asarray(slot_col, np.int32)
        nat = np.asarray(slot_nat, np.int64)
        col_slot = np.lexsort((nat, slot_col_a)).astype(np.int32)
        col_ptr = np.concatenate([[0], np.cumsum(np.bincount(slot_col_a, minlength=nb))])
        degree = np.diff(np.asarray(slot_ptr))
        packed = np.zeros((len(slot_col) + 32, 2), dtype=np.int32)
        packed[: len(slot_col), 0] = np.asarray(slot_col) * z
        packed[: len(slot_col), 1] = slot_shift
        return {
            "lifting": np.int32(z),
            "base_rows": np.int32(mb),
            "base_cols": np.int32(nb),
            "row_ptr": np.asarray(slot_ptr, np.int32),
            "slot_col": slot_col_a,
            "slot_shift": np.asarray(slot_shift, np.int32),
            "row_id": np.asarray(block_order, np.int32),
            "col_ptr": col_ptr.astype(np.int32),
            "col_slot": col_slot,
            "slot_pos": np.asarray(slot_pos, np.int32),
            "slot_row": np.asarray(slot_row, np.int32),
            "max_degree": np.int32(degree.max()),
            "slot_packed": packed,
        }

    @cached_property
    def _var_view(self) -> tuple[np.ndarray, np.ndarray]:
        """``(var_ptr, var_edges)``: per variable, its schedule-CSR edge ids in
        ascending natural check index (the flooding sum order)."""
        row_ptr, cols, _ = self._schedule
        order = self._schedule_rows
        deg = np.diff(row_ptr)
        nat_row = np.repeat(order, deg)  # natural check of each schedule edge
        edges = np.lexsort((nat_row, cols)).astype(np.int32)
        var_ptr = np.concatenate(
            [[0], np.cumsum(np.bincount(cols, minlength=self.num_variables))]
        ).astype(np.int32)
        return var_ptr, edges


class _Encoder(NamedTuple):
    """How a code encodes: ``kind`` in {accumulate, inverse, dense}, the
    information and parity positions and the host tables."""

    kind: str
    info: np.ndarray
    parity: np.ndarray
    tables: tuple[np.ndarray, ...]


@dataclass(frozen=True)
class MinSum:
    """Normalized (and offset) min-sum: ``R = sign * max(scale * min - offset, 0)``.

    ``schedule`` is ``"layered"`` (the reference's, ``commkit/coding/ldpc.py:
    317-349``) or ``"flooding"`` (an addition).
    """

    scale: float = 0.875
    offset: float = 0.0
    max_iterations: int = 50
    early_stop: bool = True
    schedule: str = "layered"

    def __post_init__(self) -> None:
        if not require_finite(self.scale, "scale", low=0.0) <= 1.0:
            raise ValueError(f"scale must be in (0, 1], got {self.scale}.")
        require_finite(self.offset, "offset", low=0.0, inclusive=True)
        require_int(self.max_iterations, "max_iterations", low=1)
        _check_schedule(self.schedule)


@dataclass(frozen=True)
class SumProduct:
    """Sum-product: ``R = sign * phi(sum phi(|Q|))``, ``phi(x) = -log tanh(x/2)``
    on ``|x|`` clipped to ``[1e-7, 30]`` (``commkit/coding/ldpc.py:351-371``)."""

    max_iterations: int = 50
    early_stop: bool = True
    schedule: str = "layered"

    def __post_init__(self) -> None:
        require_int(self.max_iterations, "max_iterations", low=1)
        _check_schedule(self.schedule)


def _check_schedule(schedule: str) -> None:
    if schedule not in _SCHEDULES:
        raise ValueError(f"schedule must be one of {_SCHEDULES}, got {schedule!r}.")


# -----------------------------------------------------------------------------
# Construction
# -----------------------------------------------------------------------------


def _warn_if_4_cycles(code: LDPCCode) -> None:
    """Log a warning when two bits share two checks (a cycle of length 4 in
    the Tanner graph, on which belief propagation degrades).

    Every bit contributes each pair of its checks; a pair seen at two
    different bits closes a 4-cycle.
    """
    order = np.lexsort((code.checks, code.variables))  # by bit, then check
    bits = code.variables[order].astype(np.int64)
    chks = code.checks[order].astype(np.int64)
    start = np.searchsorted(bits, np.arange(code.num_variables + 1))
    deg = np.diff(start)
    m = np.int64(code.num_checks)
    seen: list[np.ndarray] = []
    for d in np.unique(deg[deg >= 2]):
        # the bits of degree d as rows of their d checks, all pairs (i < j)
        rows = chks[start[:-1][deg == d][:, None] + np.arange(d)]
        i, j = np.triu_indices(int(d), 1)
        seen.append((rows[:, i] * m + rows[:, j]).ravel())
    if seen:
        pairs = np.concatenate(seen)
        if np.unique(pairs).size < pairs.size:
            logger.warning("%r has 4-cycles; belief propagation degrades on them.", code)


def _first_fit(ptr: np.ndarray, variables: np.ndarray, n: int) -> np.ndarray:
    """Layer of each row: the first layer none of whose earlier rows shares a bit.

    Each bit keeps the set of layers using it as a Python-integer mask.
    """
    masks = [0] * n
    cols = variables.tolist()
    bounds = ptr.tolist()
    color = np.empty(len(bounds) - 1, dtype=np.int64)
    for r in range(len(bounds) - 1):
        row = cols[bounds[r] : bounds[r + 1]]
        used = 0
        for v in row:
            used |= masks[v]
        free = ~used & (used + 1)  # lowest clear bit
        c = free.bit_length() - 1
        for v in row:
            masks[v] |= free
        color[r] = c
    return color


def _is_accumulator(m: int, p_rows: np.ndarray, p_cols: np.ndarray) -> bool:
    """Whether the parity part (entries p_rows x p_cols of an m x m block)
    is the accumulator: ones on the diagonal and just below it, nothing
    else (parity bit i in checks i and i + 1, the last one in check m - 1)."""
    if p_rows.size != 2 * m - 1:
        return False
    below = p_rows.astype(np.int64) - p_cols
    diag = below == 0
    return bool(np.all(diag | (below == 1)) and np.count_nonzero(diag) == m and
                np.unique(p_rows.astype(np.int64) * m + p_cols).size == p_rows.size)


def _build_encoder(code: LDPCCode) -> _Encoder:
    """The systematic encoder of a code: the first of three forms that
    applies (``commkit/coding/ldpc.py:445-479`` tries the same ones in the
    same order).

    - ``accumulate``: the last m bits are the parity and their part of H is
      the accumulator; information first, parity by a running XOR of the
      information syndrome.
    - ``inverse``: the parity part is invertible over GF(2) (m up to
      8192); information first, parity = Hp^-1 H_info u.
    - ``dense``: otherwise, from H row-reduced with pivots taken from the
      right: the pivot columns carry parity, the free columns the
      information, and the reduced rows give each parity bit from them.
    """
    m, n = code.num_checks, code.num_variables
    k = n - m
    parity_entry = code.variables >= k
    info_rows, info_cols = code.checks[~parity_entry], code.variables[~parity_entry]
    p_rows, p_cols = code.checks[parity_entry], code.variables[parity_entry] - k
    info, parity = np.arange(k), np.arange(k, n)
    if _is_accumulator(m, p_rows, p_cols):
        return _Encoder("accumulate", info, parity, (info_rows, info_cols))
    inv = None
    if m <= _MAX_INVERSE:
        hp = np.zeros((m, m), dtype=np.int8)
        hp[p_rows, p_cols] = 1
        try:
            inv = inverse(hp)
        except ValueError:  # singular: the dense form below
            inv = None
    if inv is not None:
        return _Encoder("inverse", info, parity, (info_rows, info_cols, inv))
    if m * n > _MAX_DENSE:
        raise ValueError(
            f"LDPCCode ({m} x {n}): the parity part is neither dual-diagonal nor "
            "invertible, and H is too large for a dense systematic generator."
        )
    h = np.zeros((m, n), dtype=np.int8)
    h[code.checks, code.variables] = 1
    reduced, pivots = row_reduce(h, reverse=True)
    free = np.setdiff1d(np.arange(n), pivots)
    # row i of the reduction: c[pivots[i]] = XOR over free f of reduced[i, f] c[f]
    return _Encoder("dense", free, pivots, (reduced[:, free].T.astype(np.int8),))


# -----------------------------------------------------------------------------
# Encoding
# -----------------------------------------------------------------------------


class CkLdpcEncoder(ctypes.Structure):
    """``ck_ldpc_encoder`` of ``include/ck_tail.h``."""

    _fields_ = [
        ("k", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("n_out", ctypes.c_int32),
        ("words", ctypes.c_int32),
        ("info_row_ptr", ctypes.c_void_p),
        ("info_cols", ctypes.c_void_p),
        ("inverse_t", ctypes.c_void_p),
    ]


def _encoder_tables(code: LDPCCode) -> tuple[np.ndarray, ...]:
    """Host tables of ``ck_ldpc_encode``: the CSR of H over the information
    columns, and Hp^-1 packed 64 columns a word, transposed to (words, m)
    (absent for a dual-diagonal Hp)."""
    enc = code._encoder
    m, k = code.num_checks, code.k
    info_rows, info_cols = enc.tables[0], enc.tables[1]  # sorted by (row, col)
    row_ptr = np.concatenate([[0], np.cumsum(np.bincount(info_rows, minlength=m))])
    words = (m + 63) // 64
    tables = [row_ptr.astype(np.int32), np.asarray(info_cols, np.int32)]
    if enc.kind == "inverse":
        inv = np.zeros((m, words * 64), dtype=np.uint8)
        inv[:, :m] = enc.tables[2]
        packed = np.packbits(inv.reshape(m, words, 64), axis=-1, bitorder="little")
        tables.append(np.ascontiguousarray(packed.view("<u8").reshape(m, words).T))
    return tuple(tables)


def _encode_hip(code: LDPCCode, bits: Any) -> Any:
    """``ck_ldpc_encode`` of ``(R, k)`` int8 device rows into ``(R, n)``."""
    import torch

    enc = code._encoder
    dev = bits.device
    tabs = device_tables(code, "enc_hip", dev, lambda: _encoder_tables(code))
    m = code.num_checks
    desc = CkLdpcEncoder(code.k, m, code.n, (m + 63) // 64, tabs[0].data_ptr(),
                         tabs[1].data_ptr(), tabs[2].data_ptr() if enc.kind == "inverse" else None)  # fmt: skip
    u = bits.to(torch.int8).contiguous()
    out = torch.empty((u.shape[0], code.n), dtype=torch.int8, device=dev)
    with torch.cuda.device(dev):
        _native.check(_native.lib().ck_ldpc_encode(ctypes.byref(desc), u.data_ptr(), u.shape[0],
                                                   out.data_ptr(), _native.stream_of(dev)))  # fmt: skip
    return out


def _encode(code: LDPCCode, bits: Any) -> Any:
    """Encode ``(R, k)`` int8 device rows into ``(R, n)`` int8 codewords:
    ``ck_ldpc_encode`` for a systematic code (invertible or dual-diagonal
    parity part), a GF(2) product with the generator otherwise."""
    import torch

    if not bits.is_cuda:
        raise ValueError("the LDPC encoder takes device rows; encode() uploads host input.")
    enc = code._encoder
    if enc.kind in ("inverse", "accumulate"):
        return _encode_hip(code, bits)
    # Dense: the systematic generator of the row-reduced H (float32 sums of
    # at most k ones are exact), information and parity at their positions.
    dev = bits.device
    (g,) = device_tables(code, "enc_dense", dev, lambda: (enc.tables[0].astype(np.float32),))
    parity = (bits.to(torch.float32) @ g).to(torch.int64) & 1
    words = torch.empty((bits.shape[0], code.num_variables), dtype=torch.int8, device=dev)
    info_t, par_t = device_tables(code, "enc_pos", dev, lambda: (enc.info, enc.parity))
    words[:, info_t] = bits.to(torch.int8)
    words[:, par_t] = parity.to(torch.int8)
    return words[:, : code.n]


# -----------------------------------------------------------------------------
# Decoding
# -----------------------------------------------------------------------------


class CkLdpcCode(ctypes.Structure):
    """``ck_ldpc_code`` of ``include/ck_tail.h``."""

    _fields_ = [
        ("num_checks", ctypes.c_int32),
        ("num_variables", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("k", ctypes.c_int32),
        ("num_edges", ctypes.c_int32),
        ("num_layers", ctypes.c_int32),
        ("max_row_degree", ctypes.c_int32),
        ("row_ptr", ctypes.c_void_p),
        ("cols", ctypes.c_void_p),
        ("layer_ptr", ctypes.c_void_p),
        ("var_ptr", ctypes.c_void_p),
        ("var_edges", ctypes.c_void_p),
        ("info", ctypes.c_void_p),
        ("lifting", ctypes.c_int32),
        ("base_rows", ctypes.c_int32),
        ("base_cols", ctypes.c_int32),
        ("num_slots", ctypes.c_int32),
        ("qc_row_ptr", ctypes.c_void_p),
        ("qc_slot_col", ctypes.c_void_p),
        ("qc_slot_shift", ctypes.c_void_p),
        ("qc_row_id", ctypes.c_void_p),
        ("qc_col_ptr", ctypes.c_void_p),
        ("qc_col_slot", ctypes.c_void_p),
        ("qc_slot_pos", ctypes.c_void_p),
        ("qc_slot_row", ctypes.c_void_p),
        ("qc_slot_packed", ctypes.c_void_p),
        ("slot_cols", ctypes.c_void_p),
    ]


class CkLdpcMethod(ctypes.Structure):
    """``ck_ldpc_method`` of ``include/ck_tail.h``."""

    _fields_ = [
        ("rule", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("scale", ctypes.c_float),
        ("offset", ctypes.c_float),
        ("max_iterations", ctypes.c_int32),
        ("early_stop", ctypes.c_int32),
        ("kernel", ctypes.c_int32),
        ("frames_per_group", ctypes.c_int32),
        ("lanes_per_row", ctypes.c_int32),
        ("state_in_lds", ctypes.c_int32),
    ]


class DeviceCode(NamedTuple):
    """A code's ``ck_ldpc_code`` on one device and the tensors it points into."""

    struct: CkLdpcCode
    keep: tuple


def device_code(code: LDPCCode, device: Any) -> DeviceCode:
    """Upload a code's decoder tables once per device (cached on the code)."""
    row_ptr, cols, layer_ptr = code._schedule
    qc = code._qc_view
    var_ptr, var_edges = code._var_view
    info = code._encoder.info
    contiguous_info = info.size == 0 or info.size == info[-1] + 1

    deg = np.diff(row_ptr)
    width = next((w for w in (8, 16, 32) if deg.max() <= w), 0)

    def build() -> tuple[np.ndarray, ...]:
        base = [row_ptr, cols, layer_ptr, var_ptr, var_edges, info.astype(np.int32)]
        if qc is not None:
            base += [qc[key] for key in ("row_ptr", "slot_col", "slot_shift", "row_id",
                                         "col_ptr", "col_slot", "slot_pos", "slot_row",
                                         "slot_packed")]  # fmt: skip
        elif width:
            # (width, m): slot i of every schedule row, -1 past the row's end
            slots = np.full((width, code.num_checks), -1, dtype=np.int32)
            row_of = np.repeat(np.arange(code.num_checks), deg)
            slots[np.arange(cols.size) - row_ptr[row_of], row_of] = cols
            base.append(slots)
        return tuple(np.ascontiguousarray(a, dtype=np.int32) for a in base)

    t = device_tables(code, "decoder", device, build)
    s = CkLdpcCode()
    s.num_checks = code.num_checks
    s.num_variables = code.num_variables
    s.n = code.n
    s.k = code.k
    s.num_edges = int(cols.size)
    s.num_layers = int(layer_ptr.size - 1)
    s.max_row_degree = int(deg.max())
    s.row_ptr, s.cols, s.layer_ptr = t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr()
    s.var_ptr, s.var_edges = t[3].data_ptr(), t[4].data_ptr()
    s.info = None if contiguous_info else t[5].data_ptr()
    if qc is not None:
        s.lifting = int(qc["lifting"])
        s.base_rows = int(qc["base_rows"])
        s.base_cols = int(qc["base_cols"])
        s.num_slots = int(qc["slot_col"].size)
        (s.qc_row_ptr, s.qc_slot_col, s.qc_slot_shift, s.qc_row_id, s.qc_col_ptr,
         s.qc_col_slot, s.qc_slot_pos, s.qc_slot_row,
         s.qc_slot_packed) = (x.data_ptr() for x in t[6:15])  # fmt: skip
    elif width:
        s.slot_cols = t[6].data_ptr()
    return DeviceCode(s, t)


def code_fingerprint(code: LDPCCode) -> str:
    """A digest of the parity-check matrix (sizes, puncturing and every edge):
    two codes share it only when they are the same code."""
    import hashlib

    h = hashlib.sha256()
    h.update(np.asarray([code.num_checks, code.num_variables, code.punctured], "<i8").tobytes())
    h.update(np.ascontiguousarray(code.checks, "<i4").tobytes())
    h.update(np.ascontiguousarray(code.variables, "<i4").tobytes())
    return h.hexdigest()[:16]


#: Tile choices per (code fingerprint, rule, schedule) written by the
#: autotuner (commkit_amd.tune, file tuned.json of the package).  A tile
#: tuned for one code never applies to another, and one the library cannot
#: run (LDS, row degree) falls back to the default tile; codes without an
#: entry use the default.
TUNED: dict[tuple[str, int, int], tuple[int, int, int, int]] = {}
_DEFAULT_TILES = (0, 0, 0, -1)


def _load_tuned() -> None:
    import json
    from pathlib import Path

    path = Path(__file__).resolve().parent.parent / "tuned.json"
    if not path.exists():
        return
    for e in json.loads(path.read_text()):
        if "fingerprint" not in e:
            continue  # an entry without the code's fingerprint applies to no code
        TUNED[(str(e["fingerprint"]), int(e["rule"]), int(e["schedule"]))] = tuple(
            int(v) for v in e["tiles"])  # type: ignore[assignment]


_load_tuned()


def _fingerprint_cached(code: LDPCCode) -> str:
    fp = code.__dict__.get("_fingerprint")
    if fp is None:
        fp = code_fingerprint(code)
        code.__dict__["_fingerprint"] = fp
    return fp


def tuned_tiles(code: LDPCCode, rule: int, schedule: int) -> tuple[int, int, int, int] | None:
    """The autotuner's tile for this exact code, rule and schedule, if any."""
    if not TUNED:
        return None
    return TUNED.get((_fingerprint_cached(code), rule, schedule))


def method_struct(code: LDPCCode, method: MinSum | SumProduct) -> CkLdpcMethod:
    sp = isinstance(method, SumProduct)
    m = CkLdpcMethod()
    m.rule = 1 if sp else 0
    m.schedule = _SCHEDULES.index(method.schedule)
    m.scale = 1.0 if sp else float(np.float32(method.scale))  # type: ignore[union-attr]
    m.offset = 0.0 if sp else float(np.float32(method.offset))  # type: ignore[union-attr]
    m.max_iterations = int(method.max_iterations)
    m.early_stop = int(bool(method.early_stop))
    kernel, frames, lanes, state = tuned_tiles(code, m.rule, m.schedule) or _DEFAULT_TILES
    m.kernel, m.frames_per_group, m.lanes_per_row, m.state_in_lds = kernel, frames, lanes, state
    return m


def _decode_symbols(code: LDPCCode, x: Any, constellation: Any, noise_var: Any,
                    method: MinSum | SumProduct, *, tiles: Any = None):
    """Demap and decode in one launch: ``x`` ``(R, n / k)`` complex64 device
    symbols of a uniform separable constellation (max-log closed form, k
    bits a symbol), ``noise_var`` a float or ``(R,)`` per frame; the same
    ``(bits, app, valid, iterations)`` as ``_decode`` of the demapper's
    LLRs, bit for bit (``ck_ldpc_decode_symbols``: the decoder's frame load
    demaps, and the LLRs never pass through HBM).  Quasi-cyclic codes, on
    the layered or the flooding kernel."""
    import torch

    from ..mapping.demap import axis_tables, closed_form

    c = constellation
    if (c._axes is None or not closed_form(c._axes, "maxlog")
            or c._axes[0].levels.size != c._axes[1].levels.size):
        raise ValueError("_decode_symbols: a uniform separable constellation with axes "
                         "alike (closed-form max-log) is required")  # fmt: skip
    k = c.bits_per_symbol
    if code.n % k:
        raise ValueError(f"_decode_symbols: n = {code.n} is not a whole number of {k}-bit symbols")
    dev = x.device
    R = int(x.shape[0])
    S = code.n // k
    if x.dtype != torch.complex64 or tuple(x.shape) != (R, S):
        raise ValueError(f"_decode_symbols: x must be ({R}, {S}) complex64")
    x = x.contiguous()
    per_row = None
    inv_s2 = 1.0
    if np.ndim(noise_var) == 0:
        inv_s2 = 1.0 / max(float(noise_var), 1e-20)
    else:
        nv = np.asarray(noise_var, dtype=np.float64).reshape(-1)
        if nv.shape[0] != R or not np.all(np.isfinite(nv)) or np.any(nv < 0):
            raise ValueError("_decode_symbols: one finite noise variance >= 0 per frame")
        per_row = torch.from_numpy(1.0 / np.maximum(nv, 1e-20)).to(dev)
    app = torch.empty((R, code.n), dtype=torch.float32, device=dev)
    bits = torch.empty((R, code.k), dtype=torch.int8, device=dev)
    iters = torch.empty(R, dtype=torch.int32, device=dev)
    valid = torch.empty(R, dtype=torch.uint8, device=dev)
    if R == 0:
        return bits, app, valid.bool(), iters
    lib = _native.lib()
    with torch.cuda.device(dev):
        dc = device_code(code, dev)
        ms = method_struct(code, method)
        if tiles is not None:
            ms.kernel, ms.frames_per_group, ms.lanes_per_row, ms.state_in_lds = tiles
        re_ax, im_ax, _keep = axis_tables(c, dev)
        per = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), 1)
        if per < 0 and tiles is None and (ms.kernel, ms.frames_per_group, ms.lanes_per_row,
                                          ms.state_in_lds) != _DEFAULT_TILES:
            ms.kernel, ms.frames_per_group, ms.lanes_per_row, ms.state_in_lds = _DEFAULT_TILES
            per = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), 1)
        if per < 0:
            raise RuntimeError("ck_ldpc_workspace_bytes failed")
        chunk = R if per == 0 else max(1, min(R, _GPU_WORKSPACE // max(per, 1)))
        ws_bytes = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), chunk)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        stream = _native.stream_of(dev)
        for r0 in range(0, R, chunk):
            rc = min(chunk, R - r0)
            _native.check(lib.ck_ldpc_decode_symbols(
                ctypes.byref(dc.struct), ctypes.byref(ms), x.data_ptr() + 8 * r0 * S,
                ctypes.byref(re_ax), ctypes.byref(im_ax), ctypes.c_double(inv_s2),
                per_row.data_ptr() + 8 * r0 if per_row is not None else None,
                app.data_ptr() + 4 * r0 * code.n, bits.data_ptr() + r0 * code.k,
                iters.data_ptr() + 4 * r0, valid.data_ptr() + r0, rc,
                ws.data_ptr() if ws_bytes else None, stream,
            ))  # fmt: skip
    return bits, app, valid.bool(), iters


def _decode(code: LDPCCode, llrs: Any, method: MinSum | SumProduct, *, tiles: Any = None):
    """Decode ``(R, n)`` float32 device rows: ``(bits, app, valid, iterations)``."""
    import torch

    dev = llrs.device
    R = int(llrs.shape[0])
    llrs = llrs.contiguous()
    app = torch.empty((R, code.n), dtype=torch.float32, device=dev)
    bits = torch.empty((R, code.k), dtype=torch.int8, device=dev)
    iters = torch.empty(R, dtype=torch.int32, device=dev)
    valid = torch.empty(R, dtype=torch.uint8, device=dev)
    if R == 0:
        return bits, app, valid.bool(), iters
    lib = _native.lib()
    with torch.cuda.device(dev):
        dc = device_code(code, dev)
        ms = method_struct(code, method)
        if tiles is not None:
            ms.kernel, ms.frames_per_group, ms.lanes_per_row, ms.state_in_lds = tiles
        per = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), 1)
        if per < 0 and tiles is None and (ms.kernel, ms.frames_per_group, ms.lanes_per_row,
                                          ms.state_in_lds) != _DEFAULT_TILES:
            # a tuned tile this device cannot run: the default tile instead
            ms.kernel, ms.frames_per_group, ms.lanes_per_row, ms.state_in_lds = _DEFAULT_TILES
            per = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), 1)
        if per < 0:
            _native.check(lib.ck_ldpc_decode(ctypes.byref(dc.struct), ctypes.byref(ms),
                                             None, None, None, None, None, 0, None, None))  # fmt: skip
            raise RuntimeError("ck_ldpc_workspace_bytes failed")
        chunk = R if per == 0 else max(1, min(R, _GPU_WORKSPACE // max(per, 1)))
        ws_bytes = lib.ck_ldpc_workspace_bytes(ctypes.byref(dc.struct), ctypes.byref(ms), chunk)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        stream = _native.stream_of(dev)
        for r0 in range(0, R, chunk):
            rc = min(chunk, R - r0)
            _native.check(lib.ck_ldpc_decode(
                ctypes.byref(dc.struct), ctypes.byref(ms),
                llrs.data_ptr() + 4 * r0 * code.n, app.data_ptr() + 4 * r0 * code.n,
                bits.data_ptr() + r0 * code.k, iters.data_ptr() + 4 * r0,
                valid.data_ptr() + r0, rc, ws.data_ptr() if ws_bytes else None, stream,
            ))  # fmt: skip
    return bits, app, valid.bool(), iters
