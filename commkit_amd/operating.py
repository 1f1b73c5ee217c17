"""Operating points of the seeded codes: the waterfall Es/N0.

The workloads place frames relative to a code's waterfall ("0.3 dB above
the waterfall", "uniform in [waterfall - 0.5, waterfall + 1] dB").  The
reference has no such quantity in Es/N0 (its nearest tools are the EXIT
threshold ``commkit/coding/exit.py:720-780``, the BICM GMI
``commkit/metrics/awgn.py:477`` and Monte-Carlo error-rate curves
``commkit/coding/performance.py:172-375``), so this module fixes one
operational definition and measures it with the library's own decoder:

    wf(code, constellation) = the lowest Es/N0 on a 0.05 dB grid at which
    at most 1 % of 1024 seeded frames fail to decode (layered normalized
    min-sum, scale 0.875, at most 50 iterations, early stop), i.e. the
    frame-error-rate 1e-2 point of the Monte-Carlo curve.

Frames are made by :func:`commkit_amd.workloads.make_frames` (data seed
``2000 + grid index``) and demapped with max-log LLRs.  The measured
values are kept in :data:`WATERFALL_DB`; :func:`calibrate` recomputes one
on a GPU (``python -m commkit_amd.operating``).
"""

from __future__ import annotations

import math
from typing import Any

__all__ = ["WATERFALL_DB", "SHAPED_WATERFALL_DB", "waterfall_db", "calibrate",
           "calibrate_shaped", "frame_error_rate", "shaped_frame_error_rate",
           "threshold_esn0_db"]

#: Measured waterfalls (dB) by (code name, constellation order), from
#: ``python -m commkit_amd.operating`` on an MI355X.
WATERFALL_DB: dict[tuple[str, int], float] = {
    ("r4/5-z360", 64): 16.65,
    ("r1/2-z27", 4): 3.35,
    ("r1/2-z27", 16): 9.2,
    ("r1/2-z27", 64): 14.15,
    ("r1/2-z27", 256): 19.0,
    ("r2/3-z96", 4): 3.75,
    ("r2/3-z96", 16): 9.65,
    ("r2/3-z96", 64): 14.85,
    ("r2/3-z96", 256): 19.8,
    ("r5/6-z360", 4): 5.45,
    ("r5/6-z360", 16): 11.75,
    ("r5/6-z360", 64): 17.25,
    ("r5/6-z360", 256): 22.65,
}

#: Measured waterfalls (dB) of probabilistic amplitude shaping over a code
#: (:func:`commkit_amd.workloads.make_pas_frames`) by (code name, QAM order,
#: Maxwell-Boltzmann nu): exact LLRs with the prior, layered sum-product of
#: at most 30 iterations and the dematcher, otherwise the definition above.
SHAPED_WATERFALL_DB: dict[tuple[str, int, float], float] = {
    ("r5/6-z360", 256, 0.05): 13.2,
}

_GRID = 0.05
_FRAMES = 1024
_TARGET_FER = 1e-2


def waterfall_db(code_name: str, order: int) -> float:
    try:
        return WATERFALL_DB[(code_name, order)]
    except KeyError:
        raise KeyError(
            f"no measured waterfall for {code_name} with {order}-QAM; run "
            "`python -m commkit_amd.operating` on a GPU"
        ) from None


def frame_error_rate(code: Any, constellation: Any, esn0_db: float, frames: int, seed: int,
                     method: Any = None) -> float:
    """FER of ``frames`` seeded frames at ``esn0_db`` (GPU)."""
    import torch

    from .coding.ldpc import MinSum, _decode
    from .mapping.demap import llr_any
    from .workloads import make_frames

    method = method or MinSum(max_iterations=50)
    fr = make_frames(code, constellation, esn0_db, frames, seed=seed)
    llr = llr_any(constellation, fr.rx, float(fr.noise_var[0]), "maxlog")
    bits, _, _, _ = _decode(code, llr.reshape(frames, code.n), method)
    wrong = (bits != fr.info).any(dim=1)
    torch.cuda.synchronize()
    return float(wrong.float().mean())


def shaped_frame_error_rate(code: Any, constellation: Any, esn0_db: float, frames: int,
                            seed: int, method: Any = None) -> float:
    """FER of ``frames`` codewords of probabilistic amplitude shaping at
    ``esn0_db`` (GPU): exact LLRs with the prior, the code's decoder, the
    dematcher; a frame fails when its payload differs or is flagged."""
    import torch

    from .coding.ldpc import SumProduct
    from .mapping.demap import llr_any
    from .workloads import make_pas_frames

    method = method or SumProduct(max_iterations=30)
    fr = make_pas_frames(code, constellation, esn0_db, frames, seed=seed)
    llr = llr_any(constellation, fr.rx, float(fr.noise_var[0]), "exact").reshape(frames, code.n)
    bits, _, valid, _ = fr.pas._decode(llr, method)
    wrong = (bits != fr.info).any(dim=1) | ~valid
    torch.cuda.synchronize()
    return float(wrong.float().mean())


def calibrate_shaped(code_name: str, order: int, nu: float, *, verbose: bool = True) -> float:
    """Measure the waterfall of probabilistic amplitude shaping over a code
    (SumProduct, 30 iterations)."""
    from .mapping.constellation import Constellation
    from .workloads import seeded_code

    code = seeded_code(code_name)
    c = Constellation.qam(order).shaped(nu=nu)
    bits = code.rate * math.log2(order)
    es = round(10 * math.log10(2**bits - 1) / _GRID) * _GRID

    def fer(e: float) -> float:
        idx = int(round(e / _GRID))
        f = shaped_frame_error_rate(code, c, e, _FRAMES, seed=3000 + idx % 100000)
        if verbose:
            print(f"  {code_name} {order}-QAM nu={nu} Es/N0 {e:6.2f} dB: FER {f:.4f}", flush=True)
        return f

    while fer(es) > _TARGET_FER:
        es += 0.25
    while fer(es - 0.5) <= _TARGET_FER:  # coarse descent (shaping moves it far)
        es -= 0.5
    while es - _GRID > -50 and fer(es - _GRID) <= _TARGET_FER:
        es -= _GRID
    return round(es, 2)


def calibrate(code_name: str, order: int, *, start_db: float | None = None,
              verbose: bool = True) -> float:
    """Measure the waterfall of a seeded code with Gray ``order``-QAM."""
    from .mapping.constellation import Constellation
    from .workloads import seeded_code

    code = seeded_code(code_name)
    c = Constellation.qam(order)
    # Start near the BICM-free AWGN capacity of the rate in bits per symbol.
    bits = code.rate * math.log2(order)
    es = 10 * math.log10(2**bits - 1) if start_db is None else start_db
    es = round(es / _GRID) * _GRID

    def fer(e: float) -> float:
        idx = int(round(e / _GRID))
        f = frame_error_rate(code, c, e, _FRAMES, seed=2000 + idx % 100000)
        if verbose:
            print(f"  {code_name} {order}-QAM Es/N0 {e:6.2f} dB: FER {f:.4f}", flush=True)
        return f

    # Coarse: rise in 0.25 dB steps until the target holds; then back off on the fine grid.
    while fer(es) > _TARGET_FER:
        es += 0.25
    while es - _GRID > -50 and fer(es - _GRID) <= _TARGET_FER:
        es -= _GRID
    return round(es, 2)


def threshold_esn0_db(code: Any, constellation: Any, *, tol: float = 1e-3) -> float:
    """The Es/N0 (dB) at which the BICM channel of ``constellation`` carries
    the code's belief-propagation threshold: where the NGMI,
    ``1 - (H(X) - GMI) / m``, equals :func:`~commkit_amd.coding.exit.
    ldpc_threshold` (bisection to ``tol`` dB, the GMI by
    :func:`commkit_amd.metrics.awgn_gmi` on the GPU).

    The threshold is a long-code, sum-product, Gaussian-approximation
    prediction (``commkit/coding/exit.py:720-780``): the measured waterfalls
    (finite codewords, min-sum, at most 50 iterations, FER 1e-2) lie above
    it, and the gap is what those choices cost.
    """
    from .coding.exit import ldpc_threshold
    from .metrics import awgn_gmi

    thr = ldpc_threshold(code)
    if math.isnan(thr):
        return math.nan
    m = constellation.bits_per_symbol
    h = constellation.entropy

    def ngmi(e: float) -> float:
        return 1.0 - (h - float(awgn_gmi(constellation, esn0_db=e))) / m

    lo, hi = -10.0, 40.0
    while hi - lo > tol:
        mid = 0.5 * (lo + hi)
        if ngmi(mid) >= thr:
            hi = mid
        else:
            lo = mid
    return hi


if __name__ == "__main__":
    import sys

    if sys.argv[1:] == ["--thresholds"]:
        from .coding.exit import ldpc_threshold
        from .mapping import Constellation
        from .workloads import seeded_code

        print("code       QAM  rate   threshold  predicted dB  measured dB  gap dB")
        for (name, order), wf in sorted(WATERFALL_DB.items()):
            code = seeded_code(name)
            thr = ldpc_threshold(code)
            pred = threshold_esn0_db(code, Constellation.qam(order))
            print(f"{name:10s} {order:4d}  {code.k / code.n:.3f}  {thr:.4f}     {pred:7.2f}"
                  f"       {wf:7.2f}     {wf - pred:5.2f}", flush=True)  # fmt: skip
        sys.exit(0)
    pairs = [("r4/5-z360", 64), ("r1/2-z27", 16)]
    if len(sys.argv) > 1:
        pairs = [tuple(a.split(":")) for a in sys.argv[1:]]  # type: ignore[misc]
    for p in pairs:
        if len(p) == 3:  # name:order:nu, probabilistic amplitude shaping
            wf = calibrate_shaped(p[0], int(p[1]), float(p[2]))
            print(f"shaped waterfall {p[0]} {p[1]}-QAM nu={p[2]}: {wf} dB", flush=True)
        else:
            print(f"waterfall {p[0]} {p[1]}-QAM: {calibrate(p[0], int(p[1]))} dB", flush=True)
