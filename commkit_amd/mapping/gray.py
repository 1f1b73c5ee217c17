"""Gray-labeled constellation points (host, float64).

``points[i]`` is the point whose MSB-first bit label is the binary value of
``i``, at unit average power for QAM and PAM (PSK lies on the unit circle).
Follows ``commkit/mapping/gray.py:31-60`` (``gray_code``), ``:99-156``
(``_gray_points``), ``:226-298`` (PSK, ASK) and ``:301-338`` (square QAM:
the high ``k/2`` bits select the in-phase level, the low ones the
quadrature level).  Cross QAM (32, 128, ...) is outside the receive tail's
formats and raises.
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
            raise ValueError(
                f"cross QAM ({order}-QAM) is not among the receive tail's formats."
            )
        pts = _unit_power(pts)
    else:
        raise ValueError(f"Unsupported modulation type: {family}")
    pts = np.asarray(pts)
    pts.setflags(write=False)
    return pts
