"""Additive white Gaussian noise on the device: :func:`apply_awgn`.

``commkit/impairments/noise.py:38-186``: exactly one of ``esn0_db`` (the
symbol energy is measured per channel and ``N0 = Es / (Es/N0)``; complex
samples get circular noise of variance ``N0``, real samples its in-phase
half ``N0 / 2``), ``snr_db`` (variance ``P / SNR`` with ``P`` the mean
sample power) or ``noise_var`` (absolute, scalar or one per channel).  The
noise is drawn by the ``ck_awgn`` kernel (Philox4x32-10, Box-Muller) from a
seed taken from ``rng``, so a seed repeats the noise on a given device.
"""

from __future__ import annotations

import logging
from typing import Any

import numpy as np

from . import _native
from .backend import dispatch, to_host_like
from .core.signal import Signal

__all__ = ["apply_awgn"]

logger = logging.getLogger("commkit_amd")


def _generator(rng: Any) -> np.random.Generator:
    return rng if isinstance(rng, np.random.Generator) else np.random.default_rng(rng)


def _awgn(x: Any, std_rows: Any, rng: Any) -> Any:
    """``x`` (rows, len) complex64 / float32 device tensor plus noise of
    per-row standard deviation (per component) ``std_rows``."""
    import torch

    seed = int(_generator(rng).integers(0, 2**63))
    out = torch.empty_like(x)
    std = torch.as_tensor(np.asarray(std_rows, np.float32), device=x.device).contiguous()
    rows, length = (1, x.numel()) if x.ndim == 1 else (int(x.shape[0]), int(x.shape[1]))
    with torch.cuda.device(x.device):
        _native.check(_native.lib().ck_awgn(x.data_ptr(), rows, length, int(x.is_complex()),
                                            std.data_ptr(), seed, 0, out.data_ptr(),
                                            _native.stream_of(x.device)))  # fmt: skip
    return out


def _symbol_count(signal: Signal | None, n: int, sps: float) -> float:
    """Symbols the record carries: the reference's count, else ``N / sps``
    (never more than ``N / sps``)."""
    periods = n / sps
    count = periods
    if signal is not None and signal.reference is not None and signal.reference.symbols is not None:
        count = float(np.shape(signal.reference.symbols)[-1])
    return min(count, periods)


def apply_awgn(samples: Any, *, esn0_db: float | None = None, snr_db: float | None = None,
               noise_var: Any = None, sps: float | None = None, per_channel: bool = True,
               rng: Any = None) -> Any:
    """Add white Gaussian noise at an Es/N0, an SNR or an absolute variance
    (``commkit/impairments/noise.py:38-186``).

    ``samples``: ``(N,)`` or ``(C, N)`` array or tensor, or a Signal (whose
    ``sps`` is used; a disagreeing ``sps`` raises).  Returns the same kind
    (a Signal with new samples; arrays on their device), complex64 or
    float32.  Raises ValueError unless exactly one of ``esn0_db``,
    ``snr_db``, ``noise_var`` is given, for ``esn0_db`` on an array without
    ``sps``, or a per-channel ``noise_var`` of the wrong length.
    """
    import dataclasses

    import torch

    name = "apply_awgn()"
    if sum(v is not None for v in (esn0_db, snr_db, noise_var)) != 1:
        raise ValueError(f"{name}: give exactly one of esn0_db, snr_db and noise_var.")
    signal = samples if isinstance(samples, Signal) else None
    if signal is not None:
        if sps is not None and float(sps) != float(signal.sps):
            raise ValueError(f"{name}: sps={sps} disagrees with the Signal's sps={signal.sps}.")
        sps = float(signal.sps)
        samples = signal.samples
    placed = dispatch(samples, "samples")
    x = placed.tensor
    if x.ndim not in (1, 2):
        raise ValueError(f"{name}: samples must be (N,) or (C, N), got {x.ndim}-D.")
    x = x.to(torch.complex64 if x.is_complex() else torch.float32).contiguous()
    rows = 1 if x.ndim == 1 else int(x.shape[0])
    n = int(x.shape[-1])
    power = (x.abs() ** 2).to(torch.float64).reshape(rows, n).mean(dim=1).cpu().numpy()
    if noise_var is not None:
        var = np.asarray(noise_var, dtype=np.float64)
        if not np.all(np.isfinite(var) & (var >= 0)):
            raise ValueError(f"{name}: noise_var must be finite and >= 0, got {noise_var!r}.")
        if var.ndim and var.shape != (rows,):
            raise ValueError(f"{name}: noise_var must be a scalar or have one entry per "
                             f"channel ({rows},), got shape {var.shape}.")  # fmt: skip
        var = np.broadcast_to(var, (rows,))
    elif snr_db is not None:
        p = power if per_channel else np.full(rows, power.mean())
        var = p / max(10.0 ** (float(snr_db) / 10.0), 1e-20)
    else:
        if sps is None:
            raise ValueError(f"{name}: esn0_db needs sps for array input.")
        es = power * (n / _symbol_count(signal, n, float(sps)))
        if not per_channel:
            es = np.full(rows, es.mean())
        n0 = es / max(10.0 ** (float(esn0_db) / 10.0), 1e-20)
        var = n0 if x.is_complex() else n0 / 2
    # complex noise splits its variance between I and Q
    std = np.sqrt(var / 2.0) if x.is_complex() else np.sqrt(var)
    out = _awgn(x, std, rng)
    out = to_host_like(out, placed.kind)
    return dataclasses.replace(signal, samples=out) if signal is not None else out
