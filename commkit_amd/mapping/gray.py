"""Gray-labeled constellation points (host, float64).

``points[i]`` is the point whose MSB-first bit label is the binary value of
``i``, at unit average power for QAM and PAM (PSK lies on the unit circle).
Follows ``commkit/mapping/gray.py:31-60`` (``gray_code``), ``:99-156``
(``_gray_points``), ``:226-298`` (PSK, ASK), ``:301-338`` (square QAM:
the high ``k/2`` bits select the in-phase level, the low ones the
quadrature level) and ``:341-425`` (cross QAM of an odd number of bits
above 3: a quasi-Gray labelling, the same points for every label).
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np

__all__ = ["gray_code", "gray_points"]


def gray_code(n: int) -> np.ndarray:
    """``gray_code(n)[i] = i ^ (i >> 1)``: consecutive entries differ in one bit."""
    i = np.arange(1 << n, dtype=np.int64)
    return i ^ (i >> 1)


def _gray_psk(order: int) -> np.ndarray:
    k = int(np.log2(order))
    gray = gray_code(k)
    if order == 2:
        pts = np.zeros(2, dtype=np.float64)
        pts[gray] = np.array([-1.0, 1.0])
        return pts
    phases = np.arange(order) * 2 * np.pi / order
    pts = np.zeros(order, dtype=np.complex128)
    pts[gray] = np.exp(1j * phases)
    return pts


def _gray_ask(order: int, unipolar: bool = False) -> np.ndarray:
    k = int(np.log2(order))
    gray = gray_code(k)
    levels = np.linspace(-order + 1, order - 1, order)
    if unipolar:
        levels = levels - levels.min()
    pts = np.zeros(order, dtype=np.float64)
    pts[gray] = levels
    return pts


def _gray_qam_square(order: int) -> np.ndarray:
    n = int(np.log2(order)) // 2
    pam = _gray_ask(1 << n)
    s = np.arange(order)
    return pam[s >> n] + 1j * pam[s & ((1 << n) - 1)]


def _gray_rank(word: np.ndarray, bits: int) -> np.ndarray:
    """Position of each Gray code word in the Gray sequence of ``bits`` bits
    (the inverse of :func:`gray_code`): the prefix XOR of the word's bits
    from the most significant down."""
    rank = np.asarray(word, dtype=np.int64).copy()
    shift = 1
    while shift < bits:
        rank ^= rank >> shift
        shift <<= 1
    return rank


def _gray_qam_cross(order: int) -> np.ndarray:
    """Cross QAM of ``k = log2(order)`` bits, k odd and at least 5.

    A label's high ``ceil(k/2)`` bits place it in a column and its low
    ``floor(k/2)`` bits in a row of a Gray-labelled rectangle of
    ``2^ceil(k/2)`` columns by ``2^floor(k/2)`` rows (each index Gray
    coded on its own).  The outer eighth of the columns on either side then
    moves, turned a quarter, onto the top (left band) and bottom (right
    band) of the central square of columns, which makes the cross: a point
    of the left band at (column c, row r) goes to column r + (W - H) / 2,
    row H + b - 1 - c; of the right band to the same column and row
    W - b - 1 - c, where W, H are the rectangle's sides and b the band's
    width.  Neighbours across a fold differ in more than one bit
    (quasi-Gray).  Points are odd integers, centred: ``2 x - (W - 1)``.
    """
    k = int(np.log2(order))
    col_bits, row_bits = (k + 1) // 2, k // 2
    W, H = 1 << col_bits, 1 << row_bits
    label = np.arange(order, dtype=np.int64)
    col = _gray_rank(label >> row_bits, col_bits)
    row = _gray_rank(label & (H - 1), row_bits)
    band = W >> 3
    left = col < band
    right = col >= W - band
    moved = left | right
    new_col = np.where(moved, row + (W - H) // 2, col)
    new_row = np.where(left, H + band - 1 - col, np.where(right, W - band - 1 - col, row))
    return (2.0 * new_col - (W - 1)) + 1j * (2.0 * new_row - (H - 1))


def _unit_power(points: np.ndarray) -> np.ndarray:
    rms = np.sqrt(np.mean(np.abs(points) ** 2))
    return points / rms


@lru_cache(maxsize=64)
def gray_points(family: str, order: int, unipolar: bool = False) -> np.ndarray:
    """Read-only Gray-labeled points of ``family`` in label order."""
    if order < 2:
        raise ValueError("Order must be at least 2 for modulation")
    k = int(np.log2(order))
    if 2**k != order:
        raise ValueError(f"Order must be power of 2 for {family}")
    if family == "psk":
        pts = _gray_psk(order)
    elif family == "pam":
        pts = _unit_power(_gray_ask(order, unipolar=unipolar))
    elif family == "qam":
        if order == 8:
            pts = (_gray_ask(4)[:, None] + 1j * _gray_ask(2)[None, :]).ravel()
        elif k % 2 == 0:
            pts = _gray_qam_square(order)
        else:
            pts = _gray_qam_cross(order)
        pts = _unit_power(pts)
    else:
        raise ValueError(f"Unsupported modulation type: {family}")
    pts = np.asarray(pts)
    pts.setflags(write=False)
    return pts
