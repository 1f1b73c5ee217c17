"""ctypes wrappers of ``libck_oracle.so`` -- TEST INFRASTRUCTURE ONLY.

NumPy in, NumPy out.  Each wrapper names the C function (and through it the
reference lines it restates, listed in ``oracle/ck_oracle.c``); the device
random stream of the AWGN kernel is restated in NumPy at the end.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libck_oracle.so")

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float

_SIGS = {
    "oracle_ldpc_layered": [_p, _i64, _i32, _p, _p, _i32, _i32, _f32, _f32, _i32, _i32, _p, _p],
    "oracle_ldpc_layered_probe": [_p, _i64, _i32, _p, _p, _i32, _i32, _i32, _p, _p],
    "oracle_ldpc_flooding": [_p, _i64, _i32, _p, _p, _i32, _p, _p, _i32, _f32, _f32, _i32, _i32,
                             _p, _p],
    "oracle_llr_points": [_p, _p, _i64, _p, _p, _p, _p, _i32, _i32, _f32, _i32, _p],
    "oracle_llr_separable": [_p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _i32, _i32, _f32,
                             _i32, _p],
    "oracle_llr_separable_maxlog": [_p, _p, _i64, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i32,
                                    _i32, _f32, _p],
    "oracle_ccdm_match": [_p, _i64, _i64, _p, _i32, _i64, _p],
    "oracle_ccdm_dematch": [_p, _i64, _i64, _p, _i32, _i64, _i64, _p, _p],
}  # fmt: skip

_LIB: ctypes.CDLL | None = None


def build() -> str:
    """Run oracle/Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = None
            fn.argtypes = args
        _LIB = h
    return _LIB


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def _i32a(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f32a(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def ldpc_layered(app: np.ndarray, row_ptr, cols, *, sum_product: bool, scale: float = 1.0,
                 offset: float = 0.0, max_iterations: int, early_stop: bool):
    """Layered decoding of ``app`` (R, N) float32 in place; returns (iters, valid)."""
    app = np.ascontiguousarray(app, dtype=np.float32)
    row_ptr, cols = _i32a(row_ptr), _i32a(cols)
    R, N = app.shape
    iters = np.empty(R, np.int32)
    valid = np.empty(R, np.uint8)
    lib().oracle_ldpc_layered(
        _ptr(app), R, N, _ptr(row_ptr), _ptr(cols), row_ptr.size - 1, int(sum_product),
        float(np.float32(scale)), float(np.float32(offset)), max_iterations, int(early_stop),
        _ptr(iters), _ptr(valid),
    )  # fmt: skip
    return app, iters, valid.astype(bool)


def ldpc_layered_probe(app: np.ndarray, row_ptr, cols, *, max_iterations: int,
                       early_stop: bool):
    """Layered sum-product as :func:`ldpc_layered` but with phi evaluated in
    float64 and rounded once to float32: a probe of how far the float32
    algorithm moves under a change of phi's rounding alone (not the
    reference's arithmetic)."""
    app = np.ascontiguousarray(app, dtype=np.float32)
    row_ptr, cols = _i32a(row_ptr), _i32a(cols)
    R, N = app.shape
    iters = np.empty(R, np.int32)
    valid = np.empty(R, np.uint8)
    lib().oracle_ldpc_layered_probe(_ptr(app), R, N, _ptr(row_ptr), _ptr(cols), row_ptr.size - 1,
                                    max_iterations, int(early_stop), _ptr(iters), _ptr(valid))  # fmt: skip
    return app, iters, valid.astype(bool)


def ldpc_flooding(app: np.ndarray, row_ptr, cols, var_ptr, var_edges, *, sum_product: bool,
                  scale: float = 1.0, offset: float = 0.0, max_iterations: int,
                  early_stop: bool):
    app = np.ascontiguousarray(app, dtype=np.float32)
    row_ptr, cols, var_ptr, var_edges = map(_i32a, (row_ptr, cols, var_ptr, var_edges))
    R, N = app.shape
    iters = np.empty(R, np.int32)
    valid = np.empty(R, np.uint8)
    lib().oracle_ldpc_flooding(
        _ptr(app), R, N, _ptr(row_ptr), _ptr(cols), row_ptr.size - 1, _ptr(var_ptr),
        _ptr(var_edges), int(sum_product), float(np.float32(scale)), float(np.float32(offset)),
        max_iterations, int(early_stop), _ptr(iters), _ptr(valid),
    )  # fmt: skip
    return app, iters, valid.astype(bool)


def half_indices(bit_labels: np.ndarray) -> np.ndarray:
    """(2, k, M/2) int32: the points whose bit b is 0 and 1."""
    k = bit_labels.shape[1]
    return np.stack(
        [np.stack([np.flatnonzero(bit_labels[:, b] == v) for b in range(k)]) for v in (0, 1)]
    ).astype(np.int32)


def llr_points(x: np.ndarray, points: np.ndarray, bit_labels: np.ndarray, log_pmf: np.ndarray,
               inv_s2: float, exact: bool) -> np.ndarray:
    xr, xi = _f32a(np.real(x).ravel()), _f32a(np.imag(x).ravel())
    M, k = bit_labels.shape
    out = np.empty((xr.size, k), np.float32)
    cr, ci, lp = _f32a(np.real(points)), _f32a(np.imag(points)), _f32a(log_pmf)
    idx = half_indices(bit_labels)
    lib().oracle_llr_points(_ptr(xr), _ptr(xi), xr.size, _ptr(cr), _ptr(ci), _ptr(lp), _ptr(idx),
                            M, k, float(np.float32(inv_s2)), int(exact), _ptr(out))  # fmt: skip
    return out.reshape(-1)


def llr_separable(x: np.ndarray, axes, inv_s2: float, exact: bool) -> np.ndarray:
    """``axes``: two (levels, labels, log_prior, bits) tuples (host tables)."""
    xr, xi = _f32a(np.real(x).ravel()), _f32a(np.imag(x).ravel())
    (lre, labre, lpre, bre), (lim, labim, lpim, bim) = axes
    L, kb = labre.shape
    out = np.empty((xr.size, 2 * kb), np.float32)
    t = [_f32a(lre), _f32a(lim), _f32a(lpre), _f32a(lpim), half_indices(labre),
         half_indices(labim), _i32a(bre), _i32a(bim)]  # fmt: skip
    lib().oracle_llr_separable(_ptr(xr), _ptr(xi), xr.size, *[_ptr(a) for a in t], L, kb,
                               float(np.float32(inv_s2)), int(exact), _ptr(out))  # fmt: skip
    return out.reshape(-1)


def llr_separable_maxlog(x: np.ndarray, axes, inv_s2: float) -> np.ndarray:
    """``axes``: two (levels, packed labels, lo, hi, pos) tuples."""
    xr, xi = _f32a(np.real(x).ravel()), _f32a(np.imag(x).ravel())
    (lre, labre, lore, hire, pre), (lim, labim, loim, hiim, pim) = axes
    L, kb = len(lre), len(pre)
    out = np.empty((xr.size, 2 * kb), np.float32)
    t = [_f32a(lre), _i32a(labre), _i32a(lore), _i32a(hire), _i32a(pre), _f32a(lim),
         _i32a(labim), _i32a(loim), _i32a(hiim), _i32a(pim)]  # fmt: skip
    lib().oracle_llr_separable_maxlog(_ptr(xr), _ptr(xi), xr.size, *[_ptr(a) for a in t], L, kb,
                                      float(np.float32(inv_s2)), _ptr(out))  # fmt: skip
    return out.reshape(-1)


def ccdm_match(bits: np.ndarray, composition: np.ndarray) -> np.ndarray:
    """``oracle_ccdm_match``: (R, k) bits to (R, n) point indices, n = sum of the
    composition."""
    bits = np.ascontiguousarray(bits, dtype=np.int8)
    comp = np.ascontiguousarray(composition, dtype=np.int64)
    if comp.size > 256:
        raise ValueError("the oracle's matcher takes at most 256 points")
    n = int(comp.sum())
    out = np.empty((bits.shape[0], n), dtype=np.int32)
    lib().oracle_ccdm_match(_ptr(bits), bits.shape[0], bits.shape[1], _ptr(comp), comp.size, n,
                            _ptr(out))  # fmt: skip
    return out


def ccdm_dematch(symbols: np.ndarray, composition: np.ndarray, k: int,
                 buffer_bits: int) -> tuple[np.ndarray, np.ndarray]:
    """``oracle_ccdm_dematch``: (R, n) point indices to (R, k) bits and (R,) validity."""
    sym = np.ascontiguousarray(symbols, dtype=np.int32)
    comp = np.ascontiguousarray(composition, dtype=np.int64)
    if comp.size > 256:
        raise ValueError("the oracle's dematcher takes at most 256 points")
    bits = np.empty((sym.shape[0], k), dtype=np.int8)
    valid = np.empty(sym.shape[0], dtype=np.uint8)
    lib().oracle_ccdm_dematch(_ptr(sym), sym.shape[0], sym.shape[1], _ptr(comp), comp.size, k,
                              buffer_bits, _ptr(bits), _ptr(valid))  # fmt: skip
    return bits, valid.astype(bool)


# -- the device random stream (NumPy restatement) ---------------------------------

_PHILOX_M = (np.uint64(0xD2511F53), np.uint64(0xCD9E8D57))
_PHILOX_W = (np.uint64(0x9E3779B9), np.uint64(0xBB67AE85))
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(counter: np.ndarray, key: tuple[int, int]) -> np.ndarray:
    """Philox4x32-10 (Salmon et al., SC 2011) of (N, 4) uint32 counters under a
    two-word key: ten rounds of two 32x32->64 products, the high halves
    mixed with the other words and the key, the key bumped by the Weyl
    constants.  Pinned by the published known-answer vectors."""
    c = [np.asarray(counter, dtype=np.uint64)[:, i].copy() for i in range(4)]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    for _ in range(10):
        p0 = _PHILOX_M[0] * c[0]
        p1 = _PHILOX_M[1] * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & _MASK32, p1 & _MASK32,
             ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & _MASK32, p0 & _MASK32]
        k0 = (k0 + _PHILOX_W[0]) & _MASK32
        k1 = (k1 + _PHILOX_W[1]) & _MASK32
    return np.stack(c, axis=1).astype(np.uint32)


def awgn_normals(seed: int, stream: int, count: int) -> np.ndarray:
    """The first ``count`` standard normals of ``ck_awgn``'s stream, in
    float64: draw g has counter (g lo, g hi, stream lo, stream hi) and key
    (seed lo, seed hi); uniforms (x >> 8 + 1/2) 2^-24; Box-Muller on the
    pairs (0, 1) and (2, 3)."""
    groups = -(-count // 4)
    g = np.arange(groups, dtype=np.uint64)
    ctr = np.stack([g & _MASK32, g >> np.uint64(32),
                    np.full(groups, stream & 0xFFFFFFFF, np.uint64),
                    np.full(groups, stream >> 32, np.uint64)], axis=1)  # fmt: skip
    x = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32)).astype(np.float64)
    u = (np.floor(x / 256.0) + 0.5) * 2.0**-24
    r0 = np.sqrt(-2.0 * np.log(u[:, 0]))
    r1 = np.sqrt(-2.0 * np.log(u[:, 2]))
    z = np.stack([r0 * np.cos(2 * np.pi * u[:, 1]), r0 * np.sin(2 * np.pi * u[:, 1]),
                  r1 * np.cos(2 * np.pi * u[:, 3]), r1 * np.sin(2 * np.pi * u[:, 3])], axis=1)  # fmt: skip
    return z.reshape(-1)[:count]
