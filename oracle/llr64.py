"""FP64 LLRs from the definition -- TEST INFRASTRUCTURE ONLY.

``LLR_b = F(half 0) - F(half 1)`` over the points whose bit ``b`` is 0 and
1, with ``m(s) = -|x - s|^2 / sigma^2 + log P(s)`` and ``F`` the max
(max-log) or log-sum-exp (exact), as stated in ``commkit/mapping/llr.py:
79-86`` and checked by brute force in ``tests/mapping/test_llr.py:363-376``.
Every metric of every point is formed in float64 from the float64
constellation; nothing is factored per axis, so this checks the per-axis
routes of the kernels, not restates them.
"""

from __future__ import annotations

import numpy as np


def llr_fp64(x: np.ndarray, points: np.ndarray, bit_labels: np.ndarray, pmf: np.ndarray | None,
             noise_var: float, method: str, *, chunk: int = 1 << 14) -> np.ndarray:
    """float64 LLRs ``(N * k,)`` of symbols ``x`` (complex; real symbols carry
    real noise, whose likelihood doubles the variance)."""
    x = np.asarray(x).ravel()
    real = not np.iscomplexobj(x)
    nv = 2.0 * noise_var if real else noise_var
    pts = np.asarray(points, dtype=np.complex128)
    M, k = bit_labels.shape
    logp = np.zeros(M) if pmf is None else np.log(np.clip(np.asarray(pmf, np.float64), 1e-40, None))
    out = np.empty((x.size, k))
    for i0 in range(0, x.size, chunk):
        xs = x[i0 : i0 + chunk].astype(np.complex128)
        metric = -np.abs(xs[:, None] - pts[None, :]) ** 2 / nv + logp[None, :]
        for b in range(k):
            zero = metric[:, bit_labels[:, b] == 0]
            one = metric[:, bit_labels[:, b] == 1]
            if method == "maxlog":
                out[i0 : i0 + chunk, b] = zero.max(axis=1) - one.max(axis=1)
            else:
                out[i0 : i0 + chunk, b] = np.logaddexp.reduce(zero, axis=1) - np.logaddexp.reduce(
                    one, axis=1
                )
    return out.reshape(-1)


def llr_tolerance_ok(got: np.ndarray, ref: np.ndarray, rel: float = 1e-5,
                     floor: float = 0.0) -> tuple[bool, np.ndarray]:
    """Whether ``|got - ref| <= rel * max(|ref|, floor)`` everywhere; also the
    mask of elements under the floor (where pure relative ``rel`` is not
    claimed)."""
    bound = rel * np.maximum(np.abs(ref), floor)
    ok = np.abs(got.astype(np.float64) - ref) <= bound
    return bool(ok.all()), np.abs(ref) < floor
