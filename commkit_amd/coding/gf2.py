"""Linear algebra over GF(2) on the host (encoder construction).

Rows are packed into uint64 words so an elimination step is one vectorized
XOR.  ``row_reduce(m, reverse=True)`` picks pivots from the last column
backwards: the pivot set is then the rightmost independent columns, a set
fixed by the matrix alone, and the reduced row echelon form over it is
unique (the role of ``commkit/coding/_gf2.py:45-88``).  ``inverse`` inverts
a square matrix or raises ``ValueError`` when it is singular
(``_gf2.py:123-134``).
"""

from __future__ import annotations

import numpy as np

__all__ = ["row_reduce", "inverse", "rank"]


def _pack(m: np.ndarray) -> np.ndarray:
    r, n = m.shape
    words = (n + 63) // 64
    by = np.packbits(m.astype(np.uint8), axis=1, bitorder="little")
    out = np.zeros((r, words * 8), dtype=np.uint8)
    out[:, : by.shape[1]] = by
    return out.view("<u8").copy()


def _unpack(p: np.ndarray, n: int) -> np.ndarray:
    by = np.ascontiguousarray(p).view(np.uint8)
    return np.unpackbits(by, axis=1, count=n, bitorder="little").astype(np.int8)


def row_reduce(m: np.ndarray, *, reverse: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """``(rref, pivots)``: the nonzero rows of the RREF (int8) and their pivot columns."""
    a = np.asarray(m)
    if a.ndim != 2:
        raise ValueError(f"expected a matrix, got shape {a.shape}.")
    r, n = a.shape
    p = _pack(a & 1)
    pivots: list[int] = []
    top = 0
    cols = range(n - 1, -1, -1) if reverse else range(n)
    for col in cols:
        if top == r:
            break
        w = col >> 6
        bit = np.uint64(1) << np.uint64(col & 63)
        hits = np.flatnonzero(p[top:, w] & bit)
        if hits.size == 0:
            continue
        piv = top + int(hits[0])
        if piv != top:
            p[[top, piv]] = p[[piv, top]]
        rows = np.flatnonzero(p[:, w] & bit)
        rows = rows[rows != top]
        if rows.size:
            p[rows] ^= p[top]
        pivots.append(col)
        top += 1
    return _unpack(p[:top], n), np.asarray(pivots, dtype=np.int64)


def rank(m: np.ndarray) -> int:
    return int(row_reduce(m)[1].size)


def inverse(m: np.ndarray) -> np.ndarray:
    """Inverse of a square binary matrix over GF(2), int8; ValueError if singular."""
    a = np.asarray(m)
    k = a.shape[0]
    if a.shape != (k, k):
        raise ValueError(f"expected a square matrix, got shape {a.shape}.")
    rref, pivots = row_reduce(np.concatenate([a & 1, np.eye(k, dtype=np.int8)], axis=1))
    if pivots.size < k or pivots[k - 1] != k - 1:
        raise ValueError("matrix is singular over GF(2).")
    return rref[:, k:]
