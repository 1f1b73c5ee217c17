"""Per-axis factoring of a constellation (host).

A constellation is *separable* when its points are a full ``S x S`` square
lattice (the same ``S`` equally spaced levels on the real and the imaginary
axis, every combination present once), every label bit is a function of one
axis's level only, each axis carries ``log2 S`` bits, and the prior (if any)
is the product of a prior per axis.  Then a bit's LLR depends on one axis
alone, which is what the demapper's per-axis kernels use: Gray square QAM,
uniform or Maxwell-Boltzmann shaped, is separable; PSK, cross QAM and
8-QAM are not.

The factoring is the one of ``commkit/mapping/_lattice.py:24-31,73-124``
(same ``Axis`` fields, levels ascending and taken from the points
themselves, marginal priors summed over the other axis, a relative 1e-9
tolerance on lattice positions and on the product prior), reached here by
indexing every point with its lattice coordinates.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

__all__ = ["Axis", "lattice_coordinates", "separable_axes"]

# Largest distance from a lattice position, in units of the level spacing,
# and relative tolerance of the product-prior check.
_TOL = 1e-9


class Axis(NamedTuple):
    """One axis of a separable constellation."""

    levels: np.ndarray  # (L,) real levels, ascending
    labels: np.ndarray  # (L, k_axis) int8 labels of this axis's bits
    pmf: np.ndarray | None  # (L,) marginal prior, None for uniform
    bits: np.ndarray  # (k_axis,) positions of these bits in the symbol label


def _snap(values: np.ndarray, origin: float, spacing: float, side: int) -> np.ndarray | None:
    """Lattice index of each value (origin + index * spacing), or None when a
    value is off the lattice or outside its ``side`` positions."""
    pos = (values - origin) / spacing
    index = np.rint(pos)
    if np.abs(pos - index).max() > _TOL or index.min() < 0 or index.max() > side - 1:
        return None
    return index.astype(np.int64)


def lattice_coordinates(points: np.ndarray) -> tuple[int, np.ndarray] | None:
    """``(S, where)`` when ``points`` form a full ``S x S`` square lattice:
    ``where[i, q]`` is the index of the point on real level ``i`` and
    imaginary level ``q`` (levels counted from the lowest); else None."""
    points = np.asarray(points)
    side = int(np.rint(np.sqrt(points.size)))
    if side < 2 or side * side != points.size or not np.iscomplexobj(points):
        return None
    re, im = points.real, points.imag
    origin = float(re.min())
    spacing = (float(re.max()) - origin) / (side - 1)
    if not spacing > 0.0:
        return None
    col = _snap(re, origin, spacing, side)
    row = _snap(im, origin, spacing, side)
    if col is None or row is None:
        return None
    where = np.full((side, side), -1, dtype=np.int64)
    where[col, row] = np.arange(points.size)
    if (where < 0).any():  # a lattice position taken twice, another left empty
        return None
    return side, where


def separable_axes(points: np.ndarray, bit_labels: np.ndarray,
                   pmf: np.ndarray | None) -> tuple[Axis, Axis] | None:
    """The two axes of a separable constellation, or None."""
    grid = lattice_coordinates(points)
    if grid is None:
        return None
    side, where = grid
    table = np.asarray(bit_labels)[where]  # (S, S, k): labels by (real, imaginary) level
    # A real-axis bit is the same along every row of the table, an
    # imaginary-axis bit the same along every column; each bit is exactly one.
    along_im = (table == table[:, :1, :]).all(axis=(0, 1))
    along_re = (table == table[:1, :, :]).all(axis=(0, 1))
    if not np.all(along_im != along_re):
        return None
    re_bits = np.flatnonzero(along_im)
    im_bits = np.flatnonzero(along_re)
    if (1 << re_bits.size) != side or (1 << im_bits.size) != side:
        return None
    re_pmf = im_pmf = None
    if pmf is not None:
        joint = np.asarray(pmf)[where]
        re_pmf, im_pmf = joint.sum(axis=1), joint.sum(axis=0)
        if not np.allclose(joint, np.outer(re_pmf, im_pmf), rtol=_TOL, atol=0.0):
            return None
    pts = np.asarray(points)
    return (
        Axis(pts[where[:, 0]].real.copy(), table[:, 0][:, re_bits], re_pmf, re_bits),
        Axis(pts[where[0, :]].imag.copy(), table[0, :][:, im_bits], im_pmf, im_bits),
    )
