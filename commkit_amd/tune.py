"""Tile autotuner of the quasi-cyclic LDPC decoder.

The QC kernels take a tile: frames per workgroup ``F``, lanes per check row
``P`` and whether the min-sum check state lives in LDS.  Every tile decodes
bit for bit the same, for min-sum and for sum-product alike (the lanes of a
row add its phi terms in ascending edge order whatever ``P`` is): APP LLRs,
hard bits, iteration counts and valid flags.  Tiles differ only in speed,
which depends on the code (lifting, row degrees, APP size against LDS) and
on the batch.  :func:`tune` times the candidate tiles on given LLRs and
rejects any tile whose outputs differ from the first one's in a single
value; ``python -m commkit_amd.tune`` tunes the seeded codes of the
benchmark workloads and writes ``tuned.json`` beside this file, which
:mod:`commkit_amd.coding.ldpc` loads at import, keyed by the code's
fingerprint, rule and schedule (a tile is a hint: one the library cannot
run for a code falls back to its default).
"""

from __future__ import annotations

import json
from pathlib import Path
from typing import Any

__all__ = ["TUNED_FILE", "candidates", "tune", "write"]

TUNED_FILE = Path(__file__).with_name("tuned.json")
_MAX_THREADS = 768  # lanes of a QC workgroup


def candidates(code: Any, method: Any) -> list[tuple[int, int, int, int]]:
    """Tiles ``(kernel=2, F, P, state_in_lds)`` worth timing for ``code``."""
    from .coding.ldpc import SumProduct

    qc = code._qc_view
    if qc is None:
        return []
    z = int(qc["lifting"])
    sp = isinstance(method, SumProduct)
    out = []
    flood_sp = sp and method.schedule == "flooding"  # one lane per row there
    for p in (1,) if flood_sp else (1, 2, 4):
        if z * p > _MAX_THREADS:
            continue
        f = 1
        while f * z * p <= _MAX_THREADS and f <= 64:
            out.append((2, f, p, 0))
            if not sp and method.schedule == "layered":
                out.append((2, f, p, 1))
            f *= 2
    return out


def _agree(out: Any, ref: Any) -> bool:
    """Two tiles agree when bits, APP LLRs, valid flags and iteration counts
    are equal element for element (for min-sum and sum-product alike)."""
    import torch

    return all(torch.equal(a, b) for a, b in zip(out, ref))


def tune(code: Any, method: Any, llrs: Any, *, rounds: int = 7,
         verbose: bool = False) -> tuple[tuple[int, int, int, int], list[dict]]:
    """Time every candidate tile on ``llrs`` (``(R, n)`` float32, GPU); the
    fastest tile and the table of timings.  Raises if two tiles disagree.

    The tiles are timed in interleaved rounds, each tile once a round in a
    rotating order, and ranked by their median: a drift of the clock over
    the run then reaches every tile alike instead of favouring the tiles
    timed in a fast stretch."""
    import torch

    from .coding.ldpc import _decode

    feasible = []
    ref = None
    for tiles in candidates(code, method):
        try:
            out = _decode(code, llrs, method, tiles=tiles)
        except RuntimeError as e:  # a tile the library rejects (LDS)
            if verbose:
                print(f"  {tiles}: {e}", flush=True)
            continue
        if ref is None:
            ref = out
        elif not _agree(out, ref):
            raise RuntimeError(f"tile {tiles} disagrees with {feasible[0]}")
        feasible.append(tiles)
    if not feasible:
        raise RuntimeError("no quasi-cyclic tile applies to this code")
    times: dict = {t: [] for t in feasible}
    for r in range(rounds):
        for j in range(len(feasible)):
            tiles = feasible[(j + r) % len(feasible)]
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            _decode(code, llrs, method, tiles=tiles)
            b.record()
            b.synchronize()
            times[tiles].append(a.elapsed_time(b))
    rows = []
    for tiles in feasible:
        t = sorted(times[tiles])
        rows.append({"tiles": list(tiles), "ms": t[len(t) // 2]})
        if verbose:
            print(f"  {tiles}: median {t[len(t) // 2]:.4f} ms (min {t[0]:.4f})", flush=True)
    best = min(rows, key=lambda r: r["ms"])
    return tuple(best["tiles"]), rows  # type: ignore[return-value]


def write(entries: list[dict], path: Path = TUNED_FILE) -> None:
    path.write_text(json.dumps(entries, indent=1) + "\n")


def _workloads():
    """(code name, QAM order, frames, method) of the benchmark workloads."""
    from .coding.ldpc import MinSum, SumProduct

    ms = MinSum(max_iterations=50)
    return [
        ("r4/5-z360", 64, 4096, ms),
        ("r5/6-z360", 64, 3330, ms),
        ("r2/3-z96", 64, 3330, ms),
        ("r1/2-z27", 64, 3330, ms),
        ("r5/6-z360", 256, 3106, SumProduct(max_iterations=30)),
        ("r1/2-z27", 16, 64, MinSum(scale=0.75, max_iterations=20, schedule="flooding")),
    ]


def main() -> None:
    from .coding.ldpc import _SCHEDULES, code_fingerprint
    from .mapping.constellation import Constellation
    from .mapping.demap import llr_any
    from .operating import waterfall_db
    from .workloads import make_frames, seeded_code

    entries = []
    for name, order, frames, method in _workloads():
        code = seeded_code(name)
        c = Constellation.qam(order)
        fr = make_frames(code, c, waterfall_db(name, order) + 0.3, frames, seed=0)
        llrs = llr_any(c, fr.rx, float(fr.noise_var[0]), "maxlog").reshape(frames, code.n)
        print(f"{name} {order}-QAM, {frames} frames, {method}", flush=True)
        best, rows = tune(code, method, llrs, verbose=True)
        sp = type(method).__name__ == "SumProduct"
        entries.append({
            "code": name, "fingerprint": code_fingerprint(code), "n": code.num_variables,
            "lifting": int(code._qc_view["lifting"]),
            "rule": int(sp), "schedule": _SCHEDULES.index(method.schedule),
            "frames": frames, "tiles": list(best), "ms": min(r["ms"] for r in rows),
        })  # fmt: skip
        print(f"  best {best}", flush=True)
    write(entries)
    print(f"wrote {TUNED_FILE}", flush=True)


if __name__ == "__main__":
    main()
