"""EXIT analysis of LDPC codes: :func:`ldpc_exit_curves`, :func:`ldpc_threshold`,
:func:`exit_chart`.

The Gaussian-approximation EXIT curves of the variable and check nodes of a
code follow from the degrees of its graph alone
(``commkit/coding/exit.py:619-780``): with ``J(σ)`` the mutual information
of a consistent Gaussian LLR ``N(σ²/2, σ²)``, ``λ_d`` / ``ρ_d`` the fractions
of edges at variable / check nodes of degree ``d``, and ``σ_ch = J⁻¹(I_ch)``
(0 for punctured bits)::

    I_E,V(I_A) = Σ_d λ_d J(sqrt((d - 1) J⁻¹(I_A)² + σ_ch²))
    I_E,C(I_A) = Σ_d ρ_d (1 - J(sqrt(d - 1) J⁻¹(1 - I_A)))

The threshold is the smallest channel information per bit at which one
iteration through both gains information everywhere below 0.99 (bisection
to 1e-6).  These are closed-form host computations on a few thousand grid
points; no device work is involved.

``J`` is evaluated by Gauss-Hermite quadrature (160 nodes in the Gaussian
variable; the integrand ``log2(1 + e^-L)`` is smooth) on a grid of σ from
0 to 20 in steps of 1/1000, and interpolated linearly, which holds ``J``
within 1e-7 of the integral (checked against adaptive quadrature in the
tests).
"""

from __future__ import annotations

import logging
import math
from dataclasses import dataclass, field
from functools import cache
from typing import Any

import numpy as np

from .._array import require_int
from .ldpc import LDPCCode

__all__ = ["ExitChart", "ExitCurve", "exit_chart", "ldpc_exit_curves", "ldpc_threshold"]

logger = logging.getLogger("commkit_amd")

#: Information at which iterations count as converged (a degree-1 variable
#: node never exceeds its channel information, so the curves of such a code
#: meet just below (1, 1)).
CONVERGED = 0.99
_TUNNEL_GRID = np.linspace(0.0, CONVERGED, 2000)
_MAX_STEPS = 10_000


@dataclass(frozen=True, eq=False)
class ExitCurve:
    """Transfer curve of one soft-in soft-out component: the information per
    bit of its extrinsic output at each a-priori information (``commkit/
    coding/exit.py:112-155``); float64 read-only host arrays."""

    a_priori: np.ndarray
    extrinsic: np.ndarray
    label: str = ""

    def __post_init__(self) -> None:
        a = np.array(self.a_priori, dtype=np.float64)
        e = np.array(self.extrinsic, dtype=np.float64)
        if a.ndim != 1 or a.shape != e.shape or a.size < 2:
            raise ValueError("ExitCurve: a_priori and extrinsic must be 1-D of one length >= 2, "
                             f"got {a.shape} and {e.shape}.")  # fmt: skip
        if not (np.all(np.isfinite(a)) and np.all(np.isfinite(e))):
            raise ValueError("ExitCurve: a_priori and extrinsic must be finite.")
        if a[0] < 0 or a[-1] > 1 or np.any(np.diff(a) <= 0):
            raise ValueError("ExitCurve: a_priori must increase within [0, 1].")
        if e.min() < -1e-9 or e.max() > 1 + 1e-9:
            raise ValueError("ExitCurve: extrinsic must lie in [0, 1].")
        e = np.clip(e, 0.0, 1.0)
        for name, arr in (("a_priori", a), ("extrinsic", e)):
            arr.setflags(write=False)
            object.__setattr__(self, name, arr)


@dataclass(frozen=True, eq=False)
class ExitChart:
    """Two curves exchanging extrinsic information, the tunnel between them
    and the decoding trajectory (``commkit/coding/exit.py:158-228``).

    ``first`` is drawn as (a_priori, extrinsic), ``second`` with its axes
    swapped.  ``converges``: one exchange gains information at every input
    up to 0.99.  ``trajectory`` (T, 2): from (0, 0), up to the first curve
    and across to the second per iteration, until 0.99 or a stall.
    ``narrowest``: where the tunnel is narrowest (open) or first closes.
    """

    first: ExitCurve
    second: ExitCurve
    converges: bool = field(init=False)
    trajectory: np.ndarray = field(init=False)
    iterations: int = field(init=False)
    narrowest: float = field(init=False)

    def __post_init__(self) -> None:
        for name in ("first", "second"):
            curve = getattr(self, name)
            if not isinstance(curve, ExitCurve):
                raise TypeError(f"ExitChart: {name} must be an ExitCurve, got "
                                f"{type(curve).__name__}.")  # fmt: skip
            if curve.a_priori[0] > 0 or curve.a_priori[-1] < CONVERGED:
                raise ValueError(f"ExitChart: the {name} curve must span a-priori information "
                                 f"from 0 to at least {CONVERGED}.")  # fmt: skip
        gain = self._exchange(_TUNNEL_GRID) - _TUNNEL_GRID
        closed = gain <= 0
        at = int(np.argmax(closed)) if closed.any() else int(np.argmin(gain))
        path = [(0.0, 0.0)]
        x = 0.0
        for _ in range(_MAX_STEPS):
            y = float(np.interp(x, self.first.a_priori, self.first.extrinsic))
            nxt = float(np.interp(y, self.second.a_priori, self.second.extrinsic))
            path += [(x, y), (nxt, y)]
            if nxt >= CONVERGED or nxt - x < 1e-6:
                break
            x = nxt
        traj = np.asarray(path)
        traj.setflags(write=False)
        object.__setattr__(self, "converges", not bool(closed.any()))
        object.__setattr__(self, "trajectory", traj)
        object.__setattr__(self, "iterations", (len(path) - 1) // 2)
        object.__setattr__(self, "narrowest", float(_TUNNEL_GRID[at]))

    def _exchange(self, x: Any) -> Any:
        y = np.interp(x, self.first.a_priori, self.first.extrinsic)
        return np.interp(y, self.second.a_priori, self.second.extrinsic)


def exit_chart(first: ExitCurve, second: ExitCurve) -> ExitChart:
    """EXIT chart of ``first`` (whose output feeds ``second``) against
    ``second`` (``commkit/coding/exit.py:231-294``)."""
    chart = ExitChart(first, second)
    logger.info("EXIT chart of %s against %s: tunnel %s at I_A = %.3f, %d iteration(s).",
                first.label or "the first curve", second.label or "the second curve",
                "open, narrowest" if chart.converges else "closed", chart.narrowest,
                chart.iterations)  # fmt: skip
    return chart


# -- J(σ) ------------------------------------------------------------------------


@cache
def _j_table() -> tuple[np.ndarray, np.ndarray]:
    """(σ, J(σ)) on [0, 20] in steps of 1/1000, cut where J stops increasing
    in float64 (J reaches 1 to rounding)."""
    t, w = np.polynomial.hermite.hermgauss(160)
    w = w / math.sqrt(math.pi)
    sigma = np.arange(20001) / 1000.0
    llr = (0.5 * sigma**2)[:, None] + (sigma[:, None] * math.sqrt(2.0)) * t[None, :]
    info = 1.0 - (np.logaddexp(0.0, -llr) @ w) / math.log(2.0)
    info[0] = 0.0
    flat = np.flatnonzero(np.diff(info) <= 0)
    stop = int(flat[0]) if flat.size else info.size - 1
    sigma, info = sigma[: stop + 1].copy(), info[: stop + 1].copy()
    sigma.setflags(write=False)
    info.setflags(write=False)
    return sigma, info


def _j(sigma: Any) -> np.ndarray:
    s, info = _j_table()
    return np.interp(sigma, s, info)


def _j_inv(info: Any) -> np.ndarray:
    s, table = _j_table()
    return np.interp(info, table, s)


# -- the node curves ---------------------------------------------------------------


def _edge_degrees(code: LDPCCode) -> tuple[tuple, tuple]:
    """Edge-perspective degree groups: variable (d - 1, has channel, weight)
    and check (d - 1, weight)."""
    vdeg = np.bincount(code.variables, minlength=code.num_variables)
    cdeg = np.bincount(code.checks, minlength=code.num_checks)
    seen = (code.variables < code.n).astype(np.int64)  # punctured bits: no channel
    keys, counts = np.unique(np.stack([vdeg[code.variables], seen]), axis=1, return_counts=True)
    edges = code.variables.size
    var = ((keys[0] - 1).astype(np.float64), keys[1].astype(bool), counts / edges)
    dc, cc = np.unique(cdeg[code.checks], return_counts=True)
    return var, ((dc - 1).astype(np.float64), cc / edges)


def _variable_curve(info: np.ndarray, sigma_ch: float, groups: tuple) -> np.ndarray:
    others, channel, weight = groups
    a2 = _j_inv(info)[..., None] ** 2
    return np.sum(weight * _j(np.sqrt(others * a2 + np.where(channel, sigma_ch**2, 0.0))),
                  axis=-1)  # fmt: skip


def _check_curve(info: np.ndarray, groups: tuple) -> np.ndarray:
    others, weight = groups
    a = _j_inv(1.0 - np.asarray(info, dtype=np.float64))[..., None]
    return np.sum(weight * (1.0 - _j(np.sqrt(others) * a)), axis=-1)


def _tunnel_open(sigma_ch: float, var: tuple, chk: tuple) -> bool:
    after = _check_curve(_variable_curve(_TUNNEL_GRID, sigma_ch, var), chk)
    return bool(np.all(after > _TUNNEL_GRID))


def _require_ldpc(code: Any, name: str) -> None:
    if not isinstance(code, LDPCCode):
        raise TypeError(f"{name}: code must be an LDPCCode, got {type(code).__name__}.")


def ldpc_exit_curves(code: LDPCCode, *, channel_mi: float,
                     num_points: int = 1001) -> tuple[ExitCurve, ExitCurve]:
    """Closed-form EXIT curves (variable nodes, check nodes) of ``code`` at
    channel information ``channel_mi`` per bit, on ``num_points`` a-priori
    points from 0 to 1 (``commkit/coding/exit.py:619-717``).  For BPSK on
    AWGN the channel information is ``J(sqrt(8 Es/N0))``; for BICM, the
    NGMI.  Raises TypeError for a non-LDPC code, ValueError for
    ``channel_mi`` outside (0, 1) or fewer than 2 points."""
    name = "ldpc_exit_curves()"
    _require_ldpc(code, name)
    mi = float(channel_mi)
    if not 0.0 < mi < 1.0:
        raise ValueError(f"{name}: channel_mi must be in (0, 1), got {mi}.")
    num_points = require_int(num_points, f"{name}: num_points", low=2)
    var, chk = _edge_degrees(code)
    grid = np.linspace(0.0, 1.0, num_points)
    return (ExitCurve(grid, _variable_curve(grid, float(_j_inv(mi)), var),
                      f"variable nodes, I_ch = {mi:.3f}"),
            ExitCurve(grid, _check_curve(grid, chk), "check nodes"))  # fmt: skip


def ldpc_threshold(code: LDPCCode) -> float:
    """Belief-propagation threshold of ``code`` in the Gaussian approximation:
    the smallest channel information per bit at which the tunnel opens, by
    bisection to 1e-6; NaN if it never opens (``commkit/coding/exit.py:
    720-780``)."""
    _require_ldpc(code, "ldpc_threshold()")
    var, chk = _edge_degrees(code)
    lo, hi = 0.0, 1.0 - 1e-9
    if not _tunnel_open(float(_j_inv(hi)), var, chk):
        logger.warning("%r: the EXIT tunnel never opens in the Gaussian approximation.", code)
        return math.nan
    while hi - lo > 1e-6:
        mid = 0.5 * (lo + hi)
        if _tunnel_open(float(_j_inv(mid)), var, chk):
            hi = mid
        else:
            lo = mid
    logger.info("Belief-propagation threshold of %r (Gaussian approximation): channel "
                "information %.4f per bit, %.4f above the rate %.4f.", code, hi,
                hi - code.k / code.n, code.k / code.n)  # fmt: skip
    return hi
