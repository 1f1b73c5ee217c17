"""Argument validation and the SISO/MIMO shape idiom (private).

The checks every public function of the receive tail runs before it
touches a device: integer and finite-real arguments, the single-precision
warning for double-precision input, and the ``(N,)`` <-> ``(1, N)`` pair
``as_2d`` / ``restore_1d``.  Messages and exception types follow
``commkit/_array.py:34-80,173-245,375-398`` of the reference so that code
written against it sees the same errors.
"""

from __future__ import annotations

import sys
import warnings
from typing import Any

import numpy as np

__all__ = ["PrecisionWarning"]


class PrecisionWarning(UserWarning):
    """Double-precision input reduced to the tail's single working precision.

    The demapper and the decoder compute in complex64 / float32 whatever
    their input (``commkit/backend.py:38``).
    """


def require_int(
    value: Any, name: str, *, low: int | None = None, high: int | None = None
) -> int:
    """``value`` as a Python ``int`` in ``[low, high]``; bools and floats raise TypeError."""
    if isinstance(value, bool) or not isinstance(value, (int, np.integer)):
        raise TypeError(f"{name} must be an integer, got {value!r}.")
    if (low is not None and value < low) or (high is not None and value > high):
        if high is None:
            bound = f">= {low}"
        elif low is None:
            bound = f"<= {high}"
        else:
            bound = f"in [{low}, {high}]"
        raise ValueError(f"{name} must be an integer {bound}, got {value!r}.")
    return int(value)


def require_finite(
    value: Any, name: str, *, low: float | None = None, inclusive: bool = False
) -> float:
    """``value`` as a finite Python ``float`` above ``low`` (at least, if ``inclusive``)."""
    scalar = not isinstance(value, bool) and isinstance(
        value, (int, float, np.integer, np.floating)
    )
    array = isinstance(value, np.ndarray) and value.ndim == 0 and value.dtype.kind in "iuf"
    if not (scalar or array):
        raise TypeError(f"{name} must be a real number, got {value!r}.")
    v = float(value)
    if not np.isfinite(v):
        raise ValueError(f"{name} must be finite, got {value!r}.")
    if low is not None and (v < low if inclusive else v <= low):
        bound = f">= {low:g}" if inclusive else f"> {low:g}"
        raise ValueError(f"{name} must be {bound}, got {value!r}.")
    return v


def warn_if_double(dtype: Any, function_name: str, working: str = "float32") -> None:
    """Warn (PrecisionWarning) when a single-precision function gets float64/complex128.

    The warning points at the first caller frame outside this package.
    """
    name = str(dtype).replace("torch.", "")
    if name not in ("float64", "complex128"):
        return
    frame: Any = sys._getframe(1)
    level = 2
    while frame is not None and frame.f_globals.get("__name__", "").startswith(
        "commkit_amd"
    ):
        frame, level = frame.f_back, level + 1
    warnings.warn(
        f"{function_name} computes in {working}: the {name} input is reduced "
        "to single precision.",
        PrecisionWarning,
        stacklevel=level,
    )


def as_2d(arr: Any, *, name: str = "array") -> tuple[Any, bool]:
    """``(N,)`` to ``(1, N)``; ``(C, N)`` unchanged.  Other ranks raise ValueError."""
    if arr.ndim == 1:
        return arr.reshape(1, -1), True
    if arr.ndim == 2:
        return arr, False
    raise ValueError(f"{name} must be 1-D (N,) or 2-D (C, N), got {arr.ndim}-D.")


def restore_1d(was_1d: bool, arr: Any) -> Any:
    """Undo :func:`as_2d` on one output."""
    return arr[0] if was_1d else arr
