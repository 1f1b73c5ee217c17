"""Small codes and helpers shared by the tests."""

from __future__ import annotations

import logging

import numpy as np

from commkit_amd.coding.ldpc import LDPCCode

# A small quasi-cyclic code (Z = 5) whose parity part (last four block
# columns) is block lower triangular with identity blocks on the diagonal,
# hence invertible: the "inverse" encoder.
SMALL_QC_BASE = np.array(
    [
        [1, 3, -1, 2, 0, -1, -1, -1],
        [-1, 0, 4, 1, 2, 0, -1, -1],
        [2, -1, 0, 3, -1, 1, 0, -1],
        [0, 4, 2, -1, 3, -1, 2, 0],
    ]
)

# A cycle-free Tanner graph (checks r0..r3 joined only through bits 1, 3, 4):
# belief propagation is exact on it.  Its parity part is singular (check 0
# has no parity bit), so it encodes through the dense systematic generator.
TREE_ROWS = [[0, 1, 3], [1, 2, 4], [3, 5, 6], [4, 7]]


def h_of(rows: list[list[int]], n: int) -> np.ndarray:
    h = np.zeros((len(rows), n), dtype=np.int8)
    for i, r in enumerate(rows):
        h[i, r] = 1
    return h


def quiet(fn, *args, **kw):
    """Call ``fn`` with the package logger's 4-cycle warnings silenced."""
    log = logging.getLogger("commkit_amd")
    level = log.level
    log.setLevel(logging.ERROR)
    try:
        return fn(*args, **kw)
    finally:
        log.setLevel(level)


def small_qc() -> LDPCCode:
    return quiet(LDPCCode.quasi_cyclic, SMALL_QC_BASE, lifting=5)


def tree_code() -> LDPCCode:
    return LDPCCode.from_parity_check(h_of(TREE_ROWS, 8))


def rows_in_schedule(code: LDPCCode) -> list[list[int]]:
    row_ptr, cols, _ = code._schedule
    return [cols[row_ptr[i] : row_ptr[i + 1]].tolist() for i in range(row_ptr.size - 1)]


def encode_host(code: LDPCCode, u: np.ndarray) -> np.ndarray:
    """Encode (R, k) bits with the library's encoder on CPU torch tensors."""
    import torch

    from commkit_amd.coding.ldpc import _encode

    return _encode(code, torch.from_numpy(np.ascontiguousarray(u, np.int8))).numpy()


def syndrome(code: LDPCCode, words: np.ndarray) -> np.ndarray:
    return (code.parity_check().astype(np.int64) @ np.asarray(words).T.astype(np.int64)) % 2
