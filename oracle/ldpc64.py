"""FP64 textbook belief propagation -- TEST INFRASTRUCTURE ONLY.

``layered_fp64`` restates ``ldpc_layered_reference`` of the reference's
tests (``tests/common/reference_impl.py:553-590``): each check, in the
given order, sends every bit the extrinsic message over the other bits'
``Q = APP - R_old`` -- min-sum ``prod sign(Q) * max(scale * min|Q| - offset, 0)``
or sum-product ``2 atanh(prod tanh(Q / 2))`` -- computed per edge by
deleting that edge (no two-minimum shortcut, no phi).

``flooding_fp64`` is the same message rule on the flooding schedule: every
check uses the previous iteration's APP; ``APP = L + sum of R``.

``tree_map`` gives the exact bitwise MAP and max-log MAP LLRs of a small
code by enumerating its codewords (``tests/coding/test_ldpc.py:219-252``):
on a cycle-free Tanner graph sum-product equals the first and min-sum with
scale 1 the second.
"""

from __future__ import annotations

from itertools import product

import numpy as np


def _message(others: np.ndarray, sum_product: bool, scale: float, offset: float) -> float:
    if sum_product:
        t = np.prod(np.tanh(others / 2))
        return 2 * np.arctanh(np.clip(t, -1 + 1e-15, 1 - 1e-15))
    mag = max(scale * np.min(np.abs(others)) - offset, 0.0)
    return float(np.prod(np.sign(others)) * mag)


def layered_fp64(llrs: np.ndarray, rows: list[list[int]], *, sum_product: bool,
                 scale: float = 1.0, offset: float = 0.0, iterations: int = 1) -> np.ndarray:
    app = np.asarray(llrs, dtype=np.float64).copy()
    msgs = [np.zeros(len(r)) for r in rows]
    for _ in range(iterations):
        for i, bits in enumerate(rows):
            q = app[bits] - msgs[i]
            new = np.array([_message(np.delete(q, e), sum_product, scale, offset)
                            for e in range(len(bits))])  # fmt: skip
            app[bits] = q + new
            msgs[i] = new
    return app


def flooding_fp64(llrs: np.ndarray, rows: list[list[int]], *, sum_product: bool,
                  scale: float = 1.0, offset: float = 0.0, iterations: int = 1) -> np.ndarray:
    L = np.asarray(llrs, dtype=np.float64).copy()
    app = L.copy()
    msgs = [np.zeros(len(r)) for r in rows]
    for _ in range(iterations):
        new_msgs = []
        for i, bits in enumerate(rows):
            q = app[bits] - msgs[i]
            new_msgs.append(np.array([_message(np.delete(q, e), sum_product, scale, offset)
                                      for e in range(len(bits))]))  # fmt: skip
        msgs = new_msgs
        app = L.copy()
        for i, bits in enumerate(rows):
            app[bits] += msgs[i]
    return app


def codewords(h: np.ndarray) -> np.ndarray:
    """All codewords of a small parity-check matrix, by enumeration."""
    n = h.shape[1]
    words = np.array(list(product([0, 1], repeat=n)), dtype=np.int64)
    ok = ((h.astype(np.int64) @ words.T) % 2 == 0).all(axis=0)
    return words[ok]


def tree_map(llrs: np.ndarray, h: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(exact MAP, max-log MAP) bit LLRs of ``llrs`` (B, n) by enumeration."""
    words = codewords(h)
    metric = -(np.asarray(llrs, np.float64) @ words.T)  # log P(y|c) up to a constant
    n = h.shape[1]
    exact = np.empty((llrs.shape[0], n))
    maxlog = np.empty_like(exact)
    for i in range(n):
        zero, one = metric[:, words[:, i] == 0], metric[:, words[:, i] == 1]
        exact[:, i] = np.logaddexp.reduce(zero, axis=1) - np.logaddexp.reduce(one, axis=1)
        maxlog[:, i] = zero.max(axis=1) - one.max(axis=1)
    return exact, maxlog
