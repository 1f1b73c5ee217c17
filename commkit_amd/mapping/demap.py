"""Array core of soft demapping: tables, routing and the HIP launches.

Three kernels in ``libck_hip.so`` compute float32 bit LLRs, one symbol per
lane, with the tables staged in LDS:

- ``ck_llr_separable_maxlog``: max-log LLRs of a uniform separable
  lattice in closed form, per axis the snapped level against the nearer of
  the two levels just outside its run of equal bits
  (``commkit/mapping/_demap.py:314-333``, ``commkit/_cuda/src/llr.cu:262-337``);
- ``ck_llr_separable``: the metrics of each axis's levels, max-log or
  exact log-sum-exp, with the per-axis log prior (``_demap.py:335-352``,
  ``llr.cu:225-259``);
- ``ck_llr_points``: the metrics of all M points (``_demap.py:150-219``,
  ``llr.cu:200-223``).

Routing follows ``Constellation.llr`` (``commkit/mapping/constellation.py:
328-332``) and ``_closed_form`` (``_demap.py:112-117``): separable and
max-log without a prior takes the closed form, separable otherwise the level
metrics, anything else the point metrics.  Sign convention: a positive LLR
means bit 0 is more likely.
"""

from __future__ import annotations

import ctypes
from typing import TYPE_CHECKING, Any

import numpy as np

from .. import _native
from ..backend import device_tables
from .constellation import pack_bits
from .lattice import Axis

if TYPE_CHECKING:
    from .constellation import Constellation

__all__ = ["LLR_MAX_K", "llr_any", "noise_variance", "opposite_levels", "log_prior"]

#: Most bits per symbol the kernels take (``commkit/_cuda/llr.py:12``).
LLR_MAX_K = 16


def noise_variance(
    noise_var: float | None,
    esn0_db: float | None,
    constellation: "Constellation",
    real: bool,
    function_name: str,
) -> float:
    """Noise variance from ``noise_var`` or ``esn0_db`` (``_demap.py:48-68``).

    ``N0 = Es / 10^(EsN0/10)`` with ``Es`` the constellation's power; real
    symbols carry the in-phase half, ``N0 / 2``.
    """
    from .._array import require_finite

    if (noise_var is None) == (esn0_db is None):
        raise ValueError(f"{function_name}: give one of noise_var and esn0_db.")
    if noise_var is not None:
        return require_finite(noise_var, f"{function_name}: noise_var", low=0.0)
    snr = require_finite(esn0_db, f"{function_name}: esn0_db")
    n0 = constellation.power() / 10 ** (snr / 10)
    return n0 / 2 if real else n0


def check_method(method: str) -> None:
    if method not in ("maxlog", "exact"):
        raise ValueError(f"Unknown method: {method}. Use 'maxlog' or 'exact'.")


def log_prior(pmf: np.ndarray | None, order: int) -> np.ndarray:
    """float32 log prior, ``log(clip(pmf, 1e-40))``; zeros for uniform (``_demap.py:125-129``)."""
    if pmf is None:
        return np.zeros(order, dtype=np.float32)
    return np.log(np.clip(np.asarray(pmf, np.float64), 1e-40, None)).astype(np.float32)


def opposite_levels(labels: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Per bit and level, the nearest level below / above with the other bit value.

    ``labels`` is ``(L, k)`` over ascending levels; returns int32 ``lo``,
    ``hi`` of shape ``(k, L)``, -1 where there is none.  These are the levels
    just outside the run of equal bits holding each level
    (``_demap.py:82-106``).
    """
    L, k = labels.shape
    lo = np.full((k, L), -1, dtype=np.int32)
    hi = np.full((k, L), -1, dtype=np.int32)
    for b in range(k):
        col = labels[:, b]
        for m in range(L):
            j = m - 1
            while j >= 0 and col[j] == col[m]:
                j -= 1
            lo[b, m] = j
            j = m + 1
            while j < L and col[j] == col[m]:
                j += 1
            hi[b, m] = j if j < L else -1
    return lo, hi


class CkAxis(ctypes.Structure):
    """``ck_axis`` of ``include/ck_tail.h``: one axis's device tables."""

    _fields_ = [
        ("levels", ctypes.c_void_p),
        ("log_prior", ctypes.c_void_p),
        ("labels", ctypes.c_void_p),
        ("lo", ctypes.c_void_p),
        ("hi", ctypes.c_void_p),
        ("pos", ctypes.c_void_p),
        ("num_levels", ctypes.c_int32),
        ("num_bits", ctypes.c_int32),
    ]


def log_prior64(pmf: np.ndarray | None, order: int) -> np.ndarray:
    """float64 log prior, ``log(clip(pmf, 1e-40))``; zeros for uniform."""
    if pmf is None:
        return np.zeros(order, dtype=np.float64)
    return np.log(np.clip(np.asarray(pmf, np.float64), 1e-40, None))


def _axis_tables(a: Axis) -> tuple[np.ndarray, ...]:
    lo, hi = opposite_levels(a.labels)
    return (
        a.levels.astype(np.float64),
        log_prior64(a.pmf, a.levels.size),
        pack_bits(a.labels).astype(np.int32),
        lo,
        hi,
        a.bits.astype(np.int32),
    )


def _ck_axis(t: tuple[Any, ...]) -> CkAxis:
    lev, lp, lab, lo, hi, pos = t
    return CkAxis(
        lev.data_ptr(), lp.data_ptr(), lab.data_ptr(), lo.data_ptr(), hi.data_ptr(),
        pos.data_ptr(), int(lev.numel()), int(pos.numel()),
    )  # fmt: skip


def axis_tables(c: "Constellation", device: Any) -> tuple[CkAxis, CkAxis, tuple]:
    """The two ``ck_axis`` structs of a separable constellation on ``device``."""
    re, im = c._axes  # type: ignore[misc]
    t = device_tables(c, "axes", device, lambda: _axis_tables(re) + _axis_tables(im))
    return _ck_axis(t[:6]), _ck_axis(t[6:]), t


def closed_form(axes: tuple[Axis, Axis], method: str) -> bool:
    """Max-log of a uniform separable constellation: the closed form."""
    return method == "maxlog" and all(a.pmf is None and a.levels.size >= 2 for a in axes)


def llr_any(c: "Constellation", x: Any, noise_var: float, method: str) -> Any:
    """LLRs of device symbols ``x`` (any shape ``(..., N)``), float32 ``(..., N*k)``."""
    import torch

    check_method(method)
    k = c.bits_per_symbol
    if k > LLR_MAX_K:
        raise ValueError(
            f"compute_llr(): GPU input supports at most {LLR_MAX_K} bits "
            f"per symbol, the constellation has {k}"
        )
    real_input = not x.is_complex()
    nv = 2.0 * noise_var if real_input else noise_var
    inv_s2 = 1.0 / max(nv, 1e-20)
    shape = x.shape
    flat = x.reshape(-1)
    flat = flat.to(torch.complex64) if flat.dtype != torch.complex64 else flat.contiguous()
    n = flat.numel()
    out = torch.empty((n, k), dtype=torch.float32, device=x.device)
    lib = _native.lib()
    stream = _native.stream_of(x.device)
    if n > 0:
        with torch.cuda.device(x.device):
            if c._axes is not None:
                re_ax, im_ax, _keep = axis_tables(c, x.device)
                if closed_form(c._axes, method):
                    _native.check(lib.ck_llr_separable_maxlog(
                        flat.data_ptr(), n, ctypes.byref(re_ax), ctypes.byref(im_ax),
                        ctypes.c_double(inv_s2), out.data_ptr(), stream,
                    ))  # fmt: skip
                else:
                    _native.check(lib.ck_llr_separable(
                        flat.data_ptr(), n, ctypes.byref(re_ax), ctypes.byref(im_ax),
                        ctypes.c_double(inv_s2), int(method == "exact"), out.data_ptr(),
                        stream,
                    ))  # fmt: skip
            else:
                pts, lp, lab = device_tables(
                    c, "points", x.device,
                    lambda: (
                        c.points.astype(np.complex128),
                        log_prior64(c.pmf, c.order),
                        pack_bits(c.bit_labels).astype(np.int32),
                    ),
                )  # fmt: skip
                _native.check(lib.ck_llr_points(
                    flat.data_ptr(), n, pts.data_ptr(), lp.data_ptr(), lab.data_ptr(),
                    c.order, k, ctypes.c_double(inv_s2), int(method == "exact"),
                    out.data_ptr(), stream,
                ))  # fmt: skip
    return out.reshape(*shape[:-1], shape[-1] * k) if len(shape) else out.reshape(-1)


def llr_permuted(c: "Constellation", x: Any, noise_var: float, method: str, perm: Any,
                 blocks: int, out: Any = None) -> Any:
    """LLRs of device symbols ``x`` written straight into ``blocks`` blocks of
    ``P = perm.numel()`` LLRs in a code's bit layout: the LLR of flat index
    ``t`` (as ``llr_any`` orders them) lands at ``[t // P, perm[t % P]]`` of
    the ``(blocks, P)`` float32 result; symbols past the blocks are not
    demapped.  Equals ``llr_any(c, x, noise_var, method)[:blocks * P]
    .view(blocks, P)[:, iperm]`` with ``iperm`` the inverse of ``perm``,
    without the gather's extra pass.  Separable constellations; ``perm`` an
    int32 device permutation (not checked)."""
    import torch

    check_method(method)
    if c._axes is None:
        raise ValueError("llr_permuted: the constellation is not separable per axis")
    if not x.is_complex():
        raise ValueError("llr_permuted: complex symbols required")
    period = int(perm.numel())
    if perm.dtype != torch.int32 or perm.device != x.device or not perm.is_contiguous():
        raise ValueError("llr_permuted: perm must be a contiguous int32 tensor on the "
                         "symbols' device")  # fmt: skip
    k = c.bits_per_symbol
    flat = x.reshape(-1).to(torch.complex64).contiguous()
    limit = int(blocks) * period
    if limit > flat.numel() * k:
        raise ValueError(f"llr_permuted: {blocks} blocks of {period} LLRs need "
                         f"{-(-limit // k)} symbols, got {flat.numel()}")  # fmt: skip
    if out is None:
        out = torch.empty((int(blocks), period), dtype=torch.float32, device=x.device)
    elif (tuple(out.shape) != (int(blocks), period) or out.dtype != torch.float32
          or not out.is_contiguous() or out.device != x.device):
        raise ValueError(f"llr_permuted: out must be a contiguous ({blocks}, {period}) "
                         "float32 tensor")  # fmt: skip
    if limit == 0:
        return out
    inv_s2 = 1.0 / max(noise_var, 1e-20)
    lib = _native.lib()
    with torch.cuda.device(x.device):
        re_ax, im_ax, _keep = axis_tables(c, x.device)
        closed = closed_form(c._axes, method)
        _native.check(lib.ck_llr_separable_permuted(
            flat.data_ptr(), flat.numel(), ctypes.byref(re_ax), ctypes.byref(im_ax),
            ctypes.c_double(inv_s2), int(closed), int(method == "exact"), out.data_ptr(),
            perm.data_ptr(), period, limit, _native.stream_of(x.device)))  # fmt: skip
    return out


def llr_rows(c: "Constellation", x: Any, noise_var: Any, method: str, out: Any = None, *,
             inv_s2: Any = None) -> Any:
    """LLRs of ``(R, S)`` complex device symbols, row ``r`` with noise variance ``noise_var[r]``.

    The rows of frames at mixed Es/N0 in one launch; row ``r`` equals
    ``llr_any(c, x[r], noise_var[r], method)`` bit for bit.  Separable
    constellations only.  ``out``: optional ``(R, S * k)`` float32 contiguous
    device tensor to write into.  ``inv_s2``: instead of ``noise_var``, the
    ``(R,)`` float64 device tensor of ``1 / max(noise_var, 1e-20)`` (no host
    transfer: the call can then be captured in a graph).  No counterpart in
    the reference (its ``compute_llr`` takes one noise variance,
    ``commkit/mapping/llr.py:24-95``).
    """
    import torch

    check_method(method)
    if c._axes is None:
        raise ValueError("llr_rows: the constellation is not separable per axis")
    if x.dim() != 2 or not x.is_complex():
        raise ValueError("llr_rows: symbols must be a (rows, symbols) complex tensor")
    R, S = int(x.shape[0]), int(x.shape[1])
    if inv_s2 is not None:
        if (inv_s2.dtype != torch.float64 or tuple(inv_s2.shape) != (R,)
                or not inv_s2.is_contiguous() or inv_s2.device != x.device):
            raise ValueError(f"llr_rows: inv_s2 must be a contiguous ({R},) float64 tensor "
                             "on the symbols' device")  # fmt: skip
        inv = inv_s2
    else:
        nv = np.asarray(noise_var, dtype=np.float64).reshape(-1)
        if nv.shape[0] != R:
            raise ValueError(f"llr_rows: {nv.shape[0]} noise variances for {R} rows")
        if not np.all(np.isfinite(nv)) or np.any(nv < 0):
            raise ValueError("llr_rows: noise variances must be finite and >= 0")
        inv = torch.from_numpy(1.0 / np.maximum(nv, 1e-20)).to(x.device)
    k = c.bits_per_symbol
    flat = x.to(torch.complex64).contiguous()
    if out is None:
        out = torch.empty((R, S * k), dtype=torch.float32, device=x.device)
    elif (out.shape != (R, S * k) or out.dtype != torch.float32 or not out.is_contiguous()
          or out.device != x.device):
        raise ValueError(f"llr_rows: out must be a contiguous ({R}, {S * k}) float32 tensor")
    if R * S == 0:
        return out
    lib = _native.lib()
    with torch.cuda.device(x.device):
        re_ax, im_ax, _keep = axis_tables(c, x.device)
        closed = closed_form(c._axes, method)
        _native.check(lib.ck_llr_separable_rows(
            flat.data_ptr(), R, S, ctypes.byref(re_ax), ctypes.byref(im_ax), inv.data_ptr(),
            int(closed), int(method == "exact"), out.data_ptr(), _native.stream_of(x.device),
        ))  # fmt: skip
    return out


class RowsDemapper:
    """Per-row-noise demapping of one separable constellation, bound once.

    For request loops that demap many small batches (a serving loop filling
    ring slots): the device tables, the ``ck_axis`` structs and the stream
    are resolved at construction, so a call is one checked C-ABI call.
    ``demapper(x, inv_s2, out)`` writes the LLRs of ``x`` (``(R, S)``
    complex64, contiguous) with row ``r``'s ``1 / sigma^2`` from ``inv_s2``
    (``(R,)`` float64, contiguous) into ``out`` (``(R, S * k)`` float32,
    contiguous), all on the constructor's device, on the stream current at
    construction; the same LLRs as :func:`llr_rows`, bit for bit.
    """

    def __init__(self, c: "Constellation", method: str, device: Any) -> None:
        import torch

        check_method(method)
        if c._axes is None:
            raise ValueError("RowsDemapper: the constellation is not separable per axis")
        self.device = torch.device(device)
        if self.device.index is None:  # "cuda": the current device
            self.device = torch.device("cuda", torch.cuda.current_device())
        with torch.cuda.device(self.device):
            self._re, self._im, self._keep = axis_tables(c, self.device)
            self._stream = _native.stream_of(self.device)
        self._closed = int(closed_form(c._axes, method))
        self._exact = int(method == "exact")
        self.k = c.bits_per_symbol
        self._lib = _native.lib()
        self._dtypes = (torch.complex64, torch.float64, torch.float32)
        self._axes = (ctypes.byref(self._re), ctypes.byref(self._im))

    def __call__(self, x: Any, inv_s2: Any, out: Any) -> Any:
        R, S = int(x.shape[0]), int(x.shape[1])
        cx, f64, f32 = self._dtypes
        if (x.dtype != cx or not x.is_contiguous() or inv_s2.dtype != f64
                or inv_s2.numel() != R or not inv_s2.is_contiguous() or out.dtype != f32
                or tuple(out.shape) != (R, S * self.k) or not out.is_contiguous()
                or x.device != self.device or inv_s2.device != self.device
                or out.device != self.device):
            raise ValueError("RowsDemapper: x (R, S) complex64, inv_s2 (R,) float64 and out "
                             "(R, S * k) float32, contiguous, on the demapper's device")  # fmt: skip
        if R * S:
            _native.check(self._lib.ck_llr_separable_rows(
                x.data_ptr(), R, S, ctypes.byref(self._re), ctypes.byref(self._im),
                inv_s2.data_ptr(), self._closed, self._exact, out.data_ptr(), self._stream))  # fmt: skip
        return out

    def raw(self, x: int, rows: int, row_len: int, inv_s2: int, out: int) -> None:
        """The call on device addresses, unchecked, for request loops that
        keep their buffers' addresses (``Slot.llrs_ptr`` of the serving
        engine): ``x`` ``(rows, row_len)`` complex64, ``inv_s2`` ``(rows,)``
        float64 and ``out`` ``(rows, row_len * k)`` float32, contiguous, on
        the demapper's device."""
        if rows * row_len:
            _native.check(self._lib.ck_llr_separable_rows(
                x, rows, row_len, self._axes[0], self._axes[1], inv_s2, self._closed,
                self._exact, out, self._stream))  # fmt: skip


class CkLlrSegment(ctypes.Structure):
    """``ck_llr_segment`` of ``include/ck_tail.h``: one request of
    ``ck_llr_separable_segments``."""

    _fields_ = [
        ("x", ctypes.c_void_p),
        ("out", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("inv_s2", ctypes.c_double),
        ("inv_rows", ctypes.c_void_p),
        ("row_len", ctypes.c_int64),
        ("perm", ctypes.c_void_p),
        ("period", ctypes.c_int64),
        ("limit", ctypes.c_int64),
    ]


class SegmentsDemapper:
    """Many requests of one separable constellation demapped in one launch
    (per 16 requests), bound once like :class:`RowsDemapper`.

    For request loops: each request is queued with :meth:`add` (its
    symbols, its output and its noise: a scalar ``noise_var`` or one
    ``1 / sigma^2`` per row; optionally a code's bit layout ``perm`` as
    :func:`llr_permuted`), and :meth:`flush` launches every queued request
    on the stream current at construction (a 16th queued request launches
    the 16 at once, so a long queue is demapped while it is still being
    filled).  Each request's LLRs equal its
    own :func:`llr_any` / :class:`RowsDemapper` / :func:`llr_permuted` call
    bit for bit.  The demapper holds the queued tensors until :meth:`flush`
    has enqueued the launch; after that, memory the caller frees is reused
    only by work ordered after the launch on the same stream (torch's
    caching allocator), so the tensors need not outlive the call.
    """

    BATCH = 16  # requests one launch takes (ck_llr_separable_segments)

    def __init__(self, c: "Constellation", method: str, device: Any) -> None:
        import torch

        check_method(method)
        if c._axes is None:
            raise ValueError("SegmentsDemapper: the constellation is not separable per axis")
        self.device = torch.device(device)
        if self.device.index is None:  # "cuda": the current device
            self.device = torch.device("cuda", torch.cuda.current_device())
        with torch.cuda.device(self.device):
            self._re, self._im, self._keep = axis_tables(c, self.device)
            self._stream = _native.stream_of(self.device)
        self._closed = int(closed_form(c._axes, method))
        self._exact = int(method == "exact")
        self.k = c.bits_per_symbol
        self._lib = _native.lib()
        self._queue: list[CkLlrSegment] = []
        self._held: list[Any] = []  # the queued requests' tensors

    def add(self, x: Any, out: Any, *, noise_var: float | None = None, inv_s2: Any = None,
            perm: Any = None) -> None:
        """Queue the LLRs of ``x`` (complex64, contiguous, ``(R, S)`` with
        ``inv_s2`` of ``(R,)`` float64, else any shape) into ``out``
        (float32, contiguous: ``x.numel() * k`` values, or with ``perm`` of
        ``P`` entries ``(blocks, P)`` holding the first ``blocks * P``
        LLRs in the code's bit layout)."""
        import torch

        if x.dtype != torch.complex64 or not x.is_contiguous() or x.device != self.device:
            raise ValueError("SegmentsDemapper.add: x must be contiguous complex64 on the "
                             "demapper's device")  # fmt: skip
        if out.dtype != torch.float32 or not out.is_contiguous() or out.device != self.device:
            raise ValueError("SegmentsDemapper.add: out must be contiguous float32 on the "
                             "demapper's device")  # fmt: skip
        n = int(x.numel())
        seg = CkLlrSegment(x.data_ptr(), out.data_ptr(), n, 1.0, None, 1, None, 1, 0)
        if (noise_var is None) == (inv_s2 is None):
            raise ValueError("SegmentsDemapper.add: give exactly one of noise_var and inv_s2")
        if inv_s2 is not None:
            R = int(x.shape[0]) if x.dim() == 2 else -1
            if (R < 1 or inv_s2.dtype != torch.float64 or inv_s2.numel() != R
                    or not inv_s2.is_contiguous() or inv_s2.device != self.device):
                raise ValueError("SegmentsDemapper.add: inv_s2 must be (R,) contiguous float64 "
                                 "for x of (R, S)")  # fmt: skip
            seg.inv_rows, seg.row_len = inv_s2.data_ptr(), int(x.shape[1])
        else:
            seg.inv_s2 = 1.0 / max(float(noise_var), 1e-20)
        if perm is not None:
            if (perm.dtype != torch.int32 or not perm.is_contiguous()
                    or perm.device != self.device or out.dim() != 2
                    or int(out.shape[1]) != perm.numel()):
                raise ValueError("SegmentsDemapper.add: perm must be contiguous int32 of P "
                                 "entries and out (blocks, P)")  # fmt: skip
            seg.perm, seg.period = perm.data_ptr(), int(perm.numel())
            seg.limit = int(out.numel())
            if seg.limit > n * self.k:
                raise ValueError(f"SegmentsDemapper.add: {out.shape[0]} blocks need "
                                 f"{-(-seg.limit // self.k)} symbols, got {n}")  # fmt: skip
        elif out.numel() != n * self.k:
            raise ValueError(f"SegmentsDemapper.add: out must hold {n * self.k} LLRs")
        self._queue.append(seg)
        self._held.append((x, out, inv_s2, perm))
        if len(self._queue) == self.BATCH:  # a full launch: start it now
            self.flush()

    def __len__(self) -> int:
        return len(self._queue)

    def flush(self) -> None:
        """Launch every queued request (one launch per 16)."""
        if not self._queue:
            return
        arr = (CkLlrSegment * len(self._queue))(*self._queue)
        self._queue = []
        try:
            _native.check(self._lib.ck_llr_separable_segments(
                arr, len(arr), ctypes.byref(self._re), ctypes.byref(self._im), self._closed,
                self._exact, self._stream))  # fmt: skip
        finally:
            self._held = []
