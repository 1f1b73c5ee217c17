"""The :class:`Constellation` value object.

Points, MSB-first bit labels and an optional probabilistic-shaping prior,
immutable and compared by value.  The surface follows
``commkit/mapping/constellation.py:55-332,438-442``: the Gray factories
``qam``/``psk``/``pam``, ``shaped(nu=|entropy=)`` (Maxwell-Boltzmann prior
``P(s) ~ exp(-nu |s'|^2)`` on the points scaled to a minimum distance of 2,
then rescaled to unit power under the prior, ``commkit/mapping/shaping.py:
284-333``), ``power``, ``map`` and ``llr``.  ``llr`` runs the HIP demapper
(see :mod:`commkit_amd.mapping.demap`).
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from functools import cached_property, lru_cache
from typing import Any

import numpy as np

from .._array import require_finite, require_int, warn_if_double
from ..backend import device_tables, dispatch, to_host_like
from .gray import gray_points
from .lattice import Axis, separable_axes

__all__ = ["Constellation", "pack_bits", "unpack_bits"]

_FAMILIES = ("qam", "psk", "pam")


def pack_bits(labels: np.ndarray) -> np.ndarray:
    """MSB-first bit rows ``(..., k)`` to int64 integers."""
    k = labels.shape[-1]
    weights = np.left_shift(1, np.arange(k - 1, -1, -1, dtype=np.int64))
    return (labels.astype(np.int64) * weights).sum(axis=-1)


def unpack_bits(indices: np.ndarray, k: int) -> np.ndarray:
    """Integers to their MSB-first ``k``-bit rows, int8 ``(N, k)``."""
    shifts = np.arange(k - 1, -1, -1, dtype=np.int64)
    return ((np.asarray(indices, np.int64)[:, None] >> shifts) & 1).astype(np.int8)


def _grid_energies(points: np.ndarray) -> np.ndarray:
    """Point energies on the grid scaled to a minimum distance of 2."""
    d = np.abs(points[:, None] - points[None, :])
    d_min = d[~np.eye(points.size, dtype=bool)].min()
    return np.abs(points * (2.0 / d_min)) ** 2


def _mb_pmf(energies: np.ndarray, nu: float) -> np.ndarray:
    log_p = -nu * energies
    p = np.exp(log_p - log_p.max())
    return p / p.sum()


def _entropy_bits(pmf: np.ndarray) -> float:
    p = pmf[pmf > 0]
    return float(-np.sum(p * np.log2(p)))


def _nu_for_entropy(energies: np.ndarray, entropy: float) -> float:
    from scipy.optimize import brentq

    min_e = float(energies.min())
    m_min = int(np.count_nonzero(np.isclose(energies, min_e, rtol=1e-9, atol=1e-12)))
    min_h, max_h = float(np.log2(m_min)), float(np.log2(energies.size))
    if not min_h < require_finite(entropy, "entropy") <= max_h:
        raise ValueError(f"entropy must be in ({min_h:g}, {max_h:g}] bits, got {entropy}.")
    if np.isclose(entropy, max_h, atol=1e-8):
        return 0.0

    def gap(nu: float) -> float:
        return _entropy_bits(_mb_pmf(energies, nu)) - entropy

    hi = 0.01
    while gap(hi) > 0:
        hi *= 10.0
    return float(brentq(gap, 0.0, hi, xtol=1e-12, rtol=1e-12))


@dataclass(frozen=True, eq=False, repr=False)
class Constellation:
    """Immutable constellation: points, bit labels and an optional prior.

    Parameters
    ----------
    points : array_like
        ``(M,)`` with ``M = 2**k``: complex (2-D) or real (PAM), distinct.
    bit_labels : array_like, optional
        ``(M, k)`` MSB-first labels, each pattern once; ``None`` gives
        natural binary (``points[i]`` carries the bits of ``i``).
    pmf : array_like, optional
        Prior of each point, non-negative, summing to 1; ``None`` = uniform.
    family : {"qam", "psk", "pam"}, optional
        Informational label set by the factories.
    """

    points: np.ndarray
    bit_labels: np.ndarray = None  # type: ignore[assignment]
    pmf: np.ndarray | None = None
    family: str | None = None

    def __post_init__(self) -> None:
        points = np.asarray(self.points)
        if points.ndim != 1 or points.size < 2:
            raise ValueError(f"points must have shape (M,) with M >= 2, got {points.shape}.")
        points = (
            points.astype(np.complex128)
            if np.iscomplexobj(points)
            else points.astype(np.float64)
        )
        if not np.all(np.isfinite(points)):
            raise ValueError("points must be finite.")
        order = points.size
        if np.unique(points).size != order:
            raise ValueError("points must be distinct.")
        k = int(np.log2(order))
        if 2**k != order:
            raise ValueError(f"The number of points must be a power of 2, got {order}.")
        if self.bit_labels is None:
            labels = unpack_bits(np.arange(order), k)
        else:
            labels = np.asarray(self.bit_labels)
            if labels.shape != (order, k):
                raise ValueError(
                    f"bit_labels must have shape ({order}, {k}), got {labels.shape}."
                )
            if not np.all((labels == 0) | (labels == 1)):
                raise ValueError("bit_labels must contain only 0 and 1.")
            labels = labels.astype(np.int8)
            if np.unique(pack_bits(labels)).size != order:
                raise ValueError("bit_labels must assign every bit pattern once.")
        pmf = None
        if self.pmf is not None:
            pmf = np.asarray(self.pmf, dtype=np.float64)
            if pmf.shape != (order,):
                raise ValueError(f"pmf must have shape ({order},), got {pmf.shape}.")
            if not np.all(np.isfinite(pmf)) or np.any(pmf < 0):
                raise ValueError("pmf must be finite and non-negative.")
            total = pmf.sum()
            if abs(total - 1.0) > 1e-6:
                raise ValueError(f"pmf must sum to 1, got {total:.8g}.")
            if abs(total - 1.0) > order * np.finfo(np.float64).eps:
                pmf = pmf / total
        if self.family is not None and self.family not in _FAMILIES:
            raise ValueError(f"family must be one of {_FAMILIES} or None.")
        for name, arr in (("points", points), ("bit_labels", labels), ("pmf", pmf)):
            if arr is not None:
                arr = np.array(arr, copy=True)
                arr.setflags(write=False)
            object.__setattr__(self, name, arr)

    # -- factories -------------------------------------------------------------

    @classmethod
    def qam(cls, order: int) -> "Constellation":
        """Gray-labeled square M-QAM (and 4x2 8-QAM) at unit average power."""
        return _named("qam", order, False)

    @classmethod
    def psk(cls, order: int) -> "Constellation":
        """Gray-labeled M-PSK on the unit circle."""
        return _named("psk", order, False)

    @classmethod
    def pam(cls, order: int, *, unipolar: bool = False) -> "Constellation":
        """Gray-labeled M-PAM (real) at unit average power."""
        return _named("pam", order, unipolar)

    def shaped(self, *, nu: float | None = None, entropy: float | None = None) -> "Constellation":
        """A copy with a Maxwell-Boltzmann prior, rescaled to unit power."""
        if (nu is None) == (entropy is None):
            raise ValueError("shaped() takes exactly one of nu or entropy.")
        energies = _grid_energies(self.points)
        if np.ptp(energies) < 1e-9:
            raise ValueError(
                "shaped() needs points of different energy; every point of "
                "this constellation has the same energy."
            )
        if entropy is not None:
            nu = _nu_for_entropy(energies, float(entropy))
        assert nu is not None
        if nu < 0:
            raise ValueError(f"nu must be >= 0, got {nu}.")
        pmf = _mb_pmf(energies, float(nu))
        scale = 1.0 / np.sqrt(np.dot(pmf, np.abs(self.points) ** 2))
        return replace(self, points=self.points * scale, pmf=pmf)

    # -- properties --------------------------------------------------------------

    @property
    def order(self) -> int:
        return int(self.points.size)

    @property
    def bits_per_symbol(self) -> int:
        return int(self.bit_labels.shape[1])

    @property
    def is_complex(self) -> bool:
        return bool(np.iscomplexobj(self.points))

    @property
    def entropy(self) -> float:
        """Entropy of the prior in bits per symbol (``log2(M)`` if uniform)."""
        return float(self.bits_per_symbol) if self.pmf is None else _entropy_bits(self.pmf)

    def power(self) -> float:
        """Average symbol power ``E[|s|^2]`` under the pmf (uniform if none)."""
        e = np.abs(self.points) ** 2
        return float(np.mean(e)) if self.pmf is None else float(np.dot(self.pmf, e))

    # -- operations ----------------------------------------------------------------

    def map(self, bits: Any) -> Any:
        """Map groups of ``k`` bits (MSB first) to symbols on the bits' device.

        ``(..., n)`` with ``n`` a multiple of ``k`` to ``(..., n / k)``
        complex64 (float32 for real points).
        """
        import torch

        placed = dispatch(bits, "bits")
        b = placed.tensor
        k = self.bits_per_symbol
        if b.shape[-1] % k:
            raise ValueError(
                f"bits: the last axis ({b.shape[-1]}) must be a multiple of "
                f"bits_per_symbol ({k})."
            )
        groups = b.reshape(*b.shape[:-1], -1, k).to(torch.int32)
        label = groups[..., 0].clone()
        for j in range(1, k):
            label <<= 1
            label |= groups[..., j]
        (pts,) = device_tables(self, "map", b.device, lambda: (self._points_by_label,))
        return to_host_like(pts[label.long()], placed.kind)

    def llr(self, symbols: Any, *, noise_var: float, method: str = "maxlog") -> Any:
        """Bit LLRs (positive: bit 0 more likely) with this prior, float32 ``(..., N*k)``.

        ``noise_var`` is ``E[|n|^2]`` of the noise on the constellation's
        scale; real symbols carry real noise, whose likelihood has twice the
        variance (``commkit/mapping/constellation.py:309-332``).
        """
        from .demap import llr_any

        nv = require_finite(noise_var, "Constellation.llr(): noise_var", low=0.0)
        placed = dispatch(symbols, "symbols")
        warn_if_double(placed.tensor.dtype, "compute_llr() / Constellation.llr()", "float32")
        out = llr_any(self, placed.tensor, nv, method)
        return to_host_like(out, placed.kind)

    # -- cached structure ------------------------------------------------------------

    @cached_property
    def _axes(self) -> tuple[Axis, Axis] | None:
        return separable_axes(self.points, self.bit_labels, self.pmf)

    @cached_property
    def _points_by_label(self) -> np.ndarray:
        dtype = np.complex64 if self.is_complex else np.float32
        pts = np.empty(self.order, dtype=dtype)
        pts[pack_bits(self.bit_labels)] = self.points
        return pts

    # -- value semantics ---------------------------------------------------------------

    def __eq__(self, other: object) -> bool:
        if self is other:
            return True
        if not isinstance(other, Constellation):
            return NotImplemented
        return (
            np.array_equal(self.points, other.points)
            and np.array_equal(self.bit_labels, other.bit_labels)
            and (
                (self.pmf is None and other.pmf is None)
                or (
                    self.pmf is not None
                    and other.pmf is not None
                    and np.array_equal(self.pmf, other.pmf)
                )
            )
        )

    def __hash__(self) -> int:
        h = self.__dict__.get("_hash")
        if h is None:
            pmf = None if self.pmf is None else (self.pmf + 0.0).tobytes()
            h = hash(((self.points + 0.0).tobytes(), self.bit_labels.tobytes(), pmf))
            self.__dict__["_hash"] = h
        return h

    def __repr__(self) -> str:
        name = self.family.upper() if self.family else "custom"
        shaped = "" if self.pmf is None else f", shaped H={self.entropy:.3f} bits"
        return f"Constellation({self.order}-{name}{shaped})"


@lru_cache(maxsize=64)
def _named(family: str, order: int, unipolar: bool) -> Constellation:
    order = require_int(order, "order", low=2)
    return Constellation(gray_points(family, order, unipolar), family=family)
