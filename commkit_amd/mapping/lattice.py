"""Square-lattice structure of a constellation (host).

``find_square_grid`` tells whether the points form a full square lattice
with the same levels on both axes, and ``find_separable_axes`` whether the
labels and the prior of that lattice factor into per-axis ones (Gray QAM,
shaped or not).  Every bit of a separable constellation depends on one axis,
so its LLR needs only that axis's levels: the demapper's fast paths.
Follows ``commkit/mapping/_lattice.py:24-31,73-124``.
"""

from __future__ import annotations

from typing import NamedTuple

import numpy as np

__all__ = ["Axis", "find_square_grid", "find_separable_axes"]


class Axis(NamedTuple):
    """One axis of a separable constellation."""

    levels: np.ndarray  # (L,) real levels, ascending
    labels: np.ndarray  # (L, k_axis) int8 labels of this axis's bits
    pmf: np.ndarray | None  # (L,) marginal prior, None for uniform
    bits: np.ndarray  # (k_axis,) positions of these bits in the symbol label


def find_square_grid(points: np.ndarray) -> tuple[float, float, int, np.ndarray] | None:
    """``(lev_min, step, side, lut)`` if ``points`` is a full square lattice.

    ``lut[i * side + q]`` is the index of the point at level ``i`` on the
    real axis and ``q`` on the imaginary one.
    """
    order = points.size
    side = int(round(order**0.5))
    if not np.iscomplexobj(points) or side * side != order or side < 2:
        return None
    re, im = points.real, points.imag
    lev_min = float(re.min())
    step = (float(re.max()) - lev_min) / (side - 1)
    if step <= 0 or abs(float(im.min()) - lev_min) > 1e-9 * step:
        return None
    gi, gq = (re - lev_min) / step, (im - lev_min) / step
    ri, rq = np.round(gi), np.round(gq)
    if max(np.abs(gi - ri).max(), np.abs(gq - rq).max()) > 1e-9:
        return None
    cell = ri.astype(np.int64) * side + rq.astype(np.int64)
    if cell.min() < 0 or cell.max() >= order or np.unique(cell).size != order:
        return None
    lut = np.empty(order, dtype=np.int64)
    lut[cell] = np.arange(order)
    lut.setflags(write=False)
    return lev_min, step, side, lut


def find_separable_axes(
    points: np.ndarray, bit_labels: np.ndarray, pmf: np.ndarray | None
) -> tuple[Axis, Axis] | None:
    """Per-axis factors of a separable square lattice, or ``None``."""
    grid = find_square_grid(points)
    if grid is None:
        return None
    _, _, side, lut = grid
    labels = bit_labels[lut].reshape(side, side, -1)  # [i, q, bit]
    on_re = np.all(labels == labels[:, :1, :], axis=(0, 1))  # constant in q
    on_im = np.all(labels == labels[:1, :, :], axis=(0, 1))  # constant in i
    if not np.all(on_re ^ on_im):
        return None
    bits_re, bits_im = np.flatnonzero(on_re), np.flatnonzero(on_im)
    if 2**bits_re.size != side or 2**bits_im.size != side:
        return None
    grid_points = points[lut].reshape(side, side)
    pmf_re = pmf_im = None
    if pmf is not None:
        p = pmf[lut].reshape(side, side)
        pmf_re, pmf_im = p.sum(axis=1), p.sum(axis=0)
        if not np.allclose(p, np.outer(pmf_re, pmf_im), rtol=1e-9, atol=0.0):
            return None
    return (
        Axis(grid_points[:, 0].real.copy(), labels[:, 0][:, bits_re], pmf_re, bits_re),
        Axis(grid_points[0, :].imag.copy(), labels[0, :][:, bits_im], pmf_im, bits_im),
    )
