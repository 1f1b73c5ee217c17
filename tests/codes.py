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
    """Encode (R, k) bits on the host in NumPy from the encoder tables the
    library builds (information syndrome, then the inverse, the accumulator
    or the dense generator); codewords (R, n) int8."""
    enc = code._encoder
    u = np.asarray(u, dtype=np.int64)
    m, k, n_mother = code.num_checks, code.k, code.num_variables
    if enc.kind == "dense":
        parity = (u @ enc.tables[0].astype(np.int64)) % 2
        words = np.zeros((u.shape[0], n_mother), dtype=np.int8)
        words[:, enc.info] = u
        words[:, enc.parity] = parity
        return words[:, : code.n]
    from scipy import sparse

    h_info = sparse.csr_matrix((np.ones(enc.tables[0].size, np.int64),
                                (enc.tables[0], enc.tables[1])), shape=(m, k))  # fmt: skip
    s = np.asarray(h_info @ u.T).T % 2
    if enc.kind == "accumulate":
        parity = np.cumsum(s, axis=1) % 2
    else:
        parity = (s @ enc.tables[2].astype(np.int64).T) % 2
    return np.concatenate([u, parity], axis=1).astype(np.int8)[:, : code.n]


def syndrome(code: LDPCCode, words: np.ndarray) -> np.ndarray:
    return (code.parity_check().astype(np.int64) @ np.asarray(words).T.astype(np.int64)) % 2
