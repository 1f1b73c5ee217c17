"""Benchmark of the soft-decision receive tail on one MI355X.

Headline workload (one GPU): the rate-4/5 QC-LDPC code drawn from seed 104
(10 x 50 base matrix, column weights 3-4, shifts in [0, 360), Z = 360,
n = 18000, k = 14400); 4096 codewords of uniform info bits, encoded, Gray
64-QAM, complex64, AWGN at 0.3 dB above the code's measured waterfall
(commkit_amd.operating).  A step is one pass of the hot path over the batch
already resident in HBM: max-log float32 LLRs of the 12,288,000 symbols,
then layered normalized min-sum (scale 0.875, at most 50 iterations,
syndrome early stop) of the 4096 frames.  The headline tail is one launch,
the demapping done in the decoder's frame load (``ck_ldpc_decode_symbols``);
the same symbols through two launches (``compute_llr``, then ``decode``)
are timed too and reported under ``other_tail``, and both tails' outputs
are checked equal.  Steps are queued back to back (no host synchronization
between them); frame latency is a step's time on the GPU.

The same JSON line carries, under ``configs``, the other workloads of the
tail, each measured in this run with its own roofline, CPU baseline and
parity check against the oracle:

- ``c1_plain``: 64 codewords of the rate-1/2 Z = 27 code, Gray 16-QAM at
  Es/N0 8 dB, max-log LLRs, flooding normalized min-sum (0.75, <= 20);
- ``c3_mixed``: 10,000 frames of three codes and four QAM orders at
  per-frame Es/N0 in [wf - 0.5, wf + 1] dB, one persistent batched decode;
- ``c4_serving``: 32 Poisson request streams (symbols in, demapped into
  the server's ring, decoded by continuous batching), p50/p99 frame latency
  at three offered loads;
- ``c5_shaped``: 2 x 2^22 shaped 256-QAM symbols (PAS codewords of the
  rate-5/6 Z = 360 code), exact LLRs with the prior, layered sum-product
  (<= 30).

Roofline: algorithmic bytes per decoder launch (per frame 4n LLRs in + 4n APP
out + k bits + 5 + I_f 8E messages) over the launch's average duration from
HIP events on the launch stream, against 8 TB/s; ``traffic`` is the HBM bytes
per launch from the FETCH_SIZE / WRITE_SIZE counters of the same kernel on the
same workload, scaled by what they read for a plain copy of known size.  CPU
baselines are the oracle (a C restatement of the reference CPU path) timed on
a bounded sample of the same workload on this host's cores.  Parity: the
oracle decodes the LLRs the GPU decoded, and bits, APP, iteration counts and
valid flags are compared frame by frame; a min-sum mismatch makes the run
exit with status 1 after printing the line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0
METRIC = "decoded info Gb/s at <=1e-5 post-FEC BER; p50/p99 frame latency"
CODE = "r4/5-z360"
ORDER = 64
FRAMES = 4096
MARGIN_DB = 0.3
LEGS = ("c1", "c3", "c4", "c5")

# HBM bytes per decoder launch from the FETCH_SIZE and WRITE_SIZE counters
# (scaled by a plain copy of known size), for the workloads they were read
# on: the headline min-sum launch and the shaped sum-product launch.
# Counters cannot be read inside this process; --traffic reads them live in
# child processes under rocprofv3 instead.
PROFILED_TRAFFIC = {
    "c2": {"frames": FRAMES, "bytes": 2.128e9, "demap_bytes": 0.3933e9,
           "profile": "profiles/ldpc_qc_layered_r45-z360_64qam_4096cw.md"},
    "c2_fused": {"frames": FRAMES, "bytes": 1.844e9, "demap_bytes": None,
                 "profile": "profiles/ldpc_qc_layered_r45-z360_64qam_4096cw.md"},
    "c5": {"frames": 3106, "bytes": 8.413e9, "demap_bytes": 0.3366e9,
           "profile": "profiles/ldpc_sum_product_shaped256_3106cw.md"},
}  # fmt: skip

# What the counters of the decoder kernels show limits them (the profiles
# above): VALU issue and waits, not HBM bandwidth.
LIMITER = {
    "c2": "VALU issue and barrier/LDS waits (SQ_INSTS_VALU fills ~56% of SIMD cycles, "
          "SQ_WAIT_ANY 43% of wave cycles); the APP lives in LDS and the check state is "
          "compressed to 16 bytes a row, so HBM traffic is a sixth of the algorithmic bytes",
    "c5": "VALU issue, the phi evaluations first (SQ_INSTS_VALU fills ~61% of SIMD cycles, "
          "69% by SQ_ACTIVE_INST_VALU; 60 lane instructions per edge and iteration, 6.5 of "
          "them transcendental), one workgroup per CU (a frame's 86 KB APP)",
    "c2_demap": "HBM: waves wait on memory 71% of their cycles while 102 VALU instructions "
                "per symbol issue 46% of the SIMD cycles; a device copy of the same bytes "
                "runs at 5.3 TB/s",
    "c5_demap": "VALU issue (~70% of SIMD cycles, 1017 instructions per symbol: float64 level "
                "metrics, one exp2 per level)",
}  # fmt: skip


def _parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--esn0", type=float, default=None,
                    help="override the headline Es/N0 (dB); default: waterfall + 0.3 dB")
    ap.add_argument("--workload", default="all",
                    choices=("all", "single", "plain", "mixed", "shaped", "serving"),
                    help="all (default): the headline line with every other workload under "
                         "'configs'; otherwise print only that workload's line")
    ap.add_argument("--legs", type=str, default=",".join(LEGS),
                    help="with --workload all: which sub-workloads to run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the reference's own published benchmarks")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU time of the headline baseline sample (5 timed runs)")
    ap.add_argument("--cpu-seconds-sub", type=float, default=5.0,
                    help="CPU time of each sub-workload's baseline sample")
    ap.add_argument("--loads", type=str, default="0.3,0.6,0.9",
                    help="serving: offered loads as fractions of the saturated throughput")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="serving: seconds of arrivals per offered load")
    ap.add_argument("--max-inflight", type=int, default=2,
                    help="serving: launches in flight (one workspace and stream each)")
    ap.add_argument("--max-launch-frames", type=int, default=0,
                    help="serving: frames per launch at most (0: the server's latency cap)")
    ap.add_argument("--bound-load", type=float, default=0.6,
                    help="serving: offered load (fraction of saturated) the p99 bound is stated at")
    ap.add_argument("--p99-bound-ms", type=float, default=10.0,
                    help="serving: stated p99 frame-latency bound at --bound-load")
    ap.add_argument("--demap-per-turn", action="store_true",
                    help="serving: the requests admitted in a loop turn demapped together "
                         "(SegmentsDemapper, up to 16 requests a launch) instead of one "
                         "demapper launch per request")
    ap.add_argument("--tail", choices=("fused", "two"), default="fused",
                    help="headline tail: demapping fused into the decoder's frame load "
                         "(one launch), or the demapping kernel then the decoder; the other "
                         "is timed too and reported under other_tail")
    ap.add_argument("--one-tail", action="store_true",
                    help="time only the headline tail (profiling runs: one kernel mix)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device-to-device twice (counter calibration) and exit")
    return ap.parse_args()


# -----------------------------------------------------------------------------
# Common helpers
# -----------------------------------------------------------------------------


def progress(**kw) -> None:
    print(json.dumps({"progress": kw}), file=sys.stderr, flush=True)


def host_cpu() -> tuple[str, int]:
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return model, threads


def time_cpu(run, target_s: float, runs: int = 5) -> tuple[float, int]:
    """Median seconds of ``runs`` timed calls of ``run(reps)`` after a warm-up,
    ``reps`` chosen so that the timed calls take about ``target_s`` in all."""
    t0 = time.perf_counter()
    run(1)
    one = max(time.perf_counter() - t0, 1e-4)
    reps = max(1, int(target_s / runs / one))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        run(reps)
        times.append(time.perf_counter() - t0)
    return float(np.median(times)) / reps, reps


def cpu_line(value_gbps: float, sample: str) -> dict:
    model, threads = host_cpu()
    return {"value": round(value_gbps, 5), "unit": "Gb/s", "cores": threads, "kind": "port",
            "sample": f"{sample}; C/OpenMP oracle, {threads} threads of {os.cpu_count()} on "
                      f"{model}, median of 5 runs"}  # fmt: skip


def decoder_bytes(code, iters: np.ndarray) -> float:
    """Algorithmic HBM bytes of one decode launch: per frame 4n (LLRs in) + 4n
    (APP out) + k (bits) + 5 (iterations, valid) + I_f * 8E (R read + write)."""
    E = code.checks.size
    per = 8 * code.n + code.k + 5
    return float(iters.size * per + 8.0 * E * iters.astype(np.float64).sum())


def roofline(nbytes: float, ms: float, kernel: str, *, traffic=None, source=None,
             limiter=None) -> dict:
    achieved = nbytes / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
           "bytes_per_launch": nbytes, "launch_ms": round(ms, 4)}  # fmt: skip
    if traffic is not None:
        out["traffic_frac"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        out["traffic_source"] = source
        if achieved > HBM_PEAK_GBS:
            out["note"] = ("frac above 1 by construction: the algorithmic bytes count a "
                           "float32 message per edge read and written in HBM every iteration, "
                           "which this kernel keeps compressed or in LDS; traffic_frac is the "
                           "HBM share the counters measure")
    if limiter is not None:
        out["limiter"] = limiter
    return out


def oracle_mod():
    sys.path.insert(0, ROOT)
    from oracle import cpu

    return cpu


def oracle_tables(c):
    """Host tables of the oracle demappers: closed form (levels, packed
    labels, opposite levels, bit positions) and level metrics (levels,
    labels, log prior, bit positions), per axis."""
    from commkit_amd.mapping.constellation import pack_bits
    from commkit_amd.mapping.demap import log_prior, opposite_levels

    closed, levels = [], []
    for a in c._axes:
        lo, hi = opposite_levels(a.labels)
        closed.append((a.levels, pack_bits(a.labels), lo, hi, a.bits))
        levels.append((a.levels, a.labels, log_prior(a.pmf, a.levels.size), a.bits))
    return closed, levels


def oracle_demap(c, x: np.ndarray, noise_var: float, method: str, tables=None) -> np.ndarray:
    cpu = oracle_mod()
    closed, levels = tables or oracle_tables(c)
    inv = float(np.float32(1.0 / noise_var))
    if method == "maxlog" and c._axes[0].pmf is None:
        return cpu.llr_separable_maxlog(x, closed, inv)
    return cpu.llr_separable(x, levels, inv, method == "exact")


def oracle_decode(code, llrs: np.ndarray, method) -> tuple:
    """(bits, app, valid, iterations) of (R, n) float32 host LLRs."""
    from commkit_amd import SumProduct

    cpu = oracle_mod()
    sp = isinstance(method, SumProduct)
    full = np.zeros((llrs.shape[0], code.num_variables), np.float32)
    full[:, : code.n] = llrs
    row_ptr, cols, _ = code._schedule
    kw = dict(sum_product=sp, scale=getattr(method, "scale", 1.0),
              offset=getattr(method, "offset", 0.0), max_iterations=method.max_iterations,
              early_stop=method.early_stop)  # fmt: skip
    if method.schedule == "flooding":
        var_ptr, var_edges = code._var_view
        app, iters, valid = cpu.ldpc_flooding(full, row_ptr, cols, var_ptr, var_edges, **kw)
    else:
        app, iters, valid = cpu.ldpc_layered(full, row_ptr, cols, **kw)
    bits = (app[:, code._encoder.info] < 0).astype(np.int8)
    return bits, app[:, : code.n], valid, iters


def to_host(out) -> tuple:
    bits, app, valid, iters = out
    return (bits.cpu().numpy(), app.cpu().numpy(), valid.cpu().numpy().astype(bool),
            iters.cpu().numpy())  # fmt: skip


def compare(got: tuple, want: tuple, frames=None, *, exact: bool) -> dict:
    """Frame-by-frame parity of GPU against oracle outputs.  exact: bits, APP,
    valid and iterations must be equal (min-sum); otherwise frames are
    identical when hard decisions of all n bits, valid and iterations agree,
    and the APP deviation is reported (sum-product)."""
    gb, ga, gv, gi = got
    ob, oa, ov, oi = want
    if exact:
        same = (gb == ob).all(1) & (ga == oa).all(1) & (gv == ov) & (gi == oi)
    else:
        same = ((ga < 0) == (oa < 0)).all(1) & (gv == ov) & (gi == oi)
    idx = np.arange(gb.shape[0]) if frames is None else np.asarray(frames)
    out = {"frames": int(gb.shape[0]), "identical": int(same.sum()),
           "divergent": idx[~same][:32].tolist(),
           "rule": "bits, APP, valid and iterations equal" if exact else
                   "hard decisions of all n bits, valid and iterations equal"}  # fmt: skip
    if not exact and same.any():
        d = np.abs(ga[same] - oa[same])
        out["app_abs_dev_quantiles"] = {q: round(float(np.quantile(d, q)), 5)
                                        for q in (0.5, 0.99, 1.0)}  # fmt: skip
        rel = d / np.maximum(np.abs(oa[same]), 1.0)
        out["app_rel_dev_quantiles"] = {q: float(f"{np.quantile(rel, q):.3g}")
                                        for q in (0.5, 0.99, 1.0)}  # fmt: skip
    return out


def llr_parity(c, x: np.ndarray, noise_var: float, method: str, got: np.ndarray,
               floor: float = 0.0) -> dict:
    """GPU LLRs of host symbols ``x`` against the FP64 definition: the
    largest |got - ref| / max(|ref|, floor) (the bar is 1e-5)."""
    sys.path.insert(0, ROOT)
    from oracle.llr64 import llr_fp64

    ref = llr_fp64(x, c.points, c.bit_labels, c.pmf, noise_var, method)
    err = np.abs(got.astype(np.float64).ravel() - ref) / np.maximum(np.abs(ref), floor)
    err = np.where(np.abs(ref) == 0, np.abs(got.ravel()), err)
    worst = float(err.max())
    return {"symbols": int(x.size), "method": method, "max_rel_err": worst,
            "floor": floor, "bar": 1e-5, "met": worst <= 1e-5}  # fmt: skip


def copy_bandwidth(dev, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Measured device-to-device copy rate (read + write bytes per second, GB/s)."""
    import torch

    x = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    b.synchronize()
    return 2.0 * nbytes * reps / (a.elapsed_time(b) * 1e-3) / 1e9


def copy_calibration(dev) -> None:
    """A plain copy of known size (256 MiB read, 256 MiB written) for scaling
    the FETCH_SIZE / WRITE_SIZE counters of the decoder's profile."""
    import torch

    x = torch.empty(1 << 28, dtype=torch.uint8, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {x.numel()} bytes twice", flush=True)


def timed_steps(steps: int, warmup: int, stages) -> tuple[float, list, list]:
    """Runs ``warmup`` untimed then ``steps`` timed passes of ``stages`` (a
    list of callables run in order on torch's current stream), with HIP
    events around each stage.  The steps are queued back to back, as a
    receiver streams batches (the host prepares step s + 1 while the GPU
    runs step s; no host synchronization between steps), and the clock is
    synchronized on both sides.  Returns (wall seconds of the timed steps,
    per-stage mean ms, per-step ms on the GPU: from the step's first event
    to its last, the time its frames spend being processed)."""
    import torch

    for _ in range(warmup):
        for fn in stages:
            fn()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
          for _ in range(steps)]  # fmt: skip
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        ev[s][0].record(stream)
        for j, fn in enumerate(stages):
            fn()
            ev[s][j + 1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    stage_ms = [float(np.mean([e[j].elapsed_time(e[j + 1]) for e in ev]))
                for j in range(len(stages))]  # fmt: skip
    step_ms = [e[0].elapsed_time(e[-1]) for e in ev]
    return elapsed, stage_ms, step_ms


def pct(step_ms) -> dict:
    return {"p50": round(float(np.percentile(step_ms, 50)), 3),
            "p99": round(float(np.percentile(step_ms, 99)), 3)}


# -----------------------------------------------------------------------------
# The reference's own published benchmarks (context, not the metric's config)
# -----------------------------------------------------------------------------

#: DVB-S2 rate 1/2 normal frames, 64 codewords, QPSK at Es/N0 1.5 dB, layered
#: MinSum() (commkit benchmarks/bench_coding.py:37-87), median 15.745 ms on an
#: RTX 4070 Ti.
REFERENCE_BENCH_MS = 15.745


def reference_benchmark(dev, reps: int = 5) -> dict:
    """The reference's decoder benchmark, rebuilt with this library's
    generate / apply_awgn / compute_llr (seeds 0 and 1) and timed with HIP
    events on the launch stream: median of ``reps`` decodes."""
    import torch

    from commkit_amd import Constellation, LDPCCode, MinSum, compute_llr, decode, metrics
    from commkit_amd.core import generate
    from commkit_amd.impairments import apply_awgn

    code = LDPCCode.dvbs2("1/2")
    sig = generate(Constellation.qam(4), num_symbols=64 * code.n // 2, symbol_rate=1e9,
                   code=code, rng=0, device="gpu")  # fmt: skip
    llrs = compute_llr(apply_awgn(sig, esn0_db=1.5, rng=1), noise_var=10 ** (-1.5 / 10))
    method = MinSum()
    res = decode(llrs, code=code, method=method)
    stream = torch.cuda.current_stream(dev)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        res = decode(llrs, code=code, method=method)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    ms = float(np.median(times))
    fer = metrics.fer(res.bits, sig.reference.info_bits, info_bits=code.k)
    return {
        "workload": "DVB-S2 r1/2 normal (n=64800, k=32400), 64 codewords, QPSK Es/N0 1.5 dB, "
                    "layered MinSum() (the reference's benchmarks/bench_coding.py)",
        "ms": round(ms, 3),
        "gbps": round(64 * code.k / (ms * 1e-3) / 1e9, 3),
        "fer": fer,
        "mean_iterations": float(res.iterations.float().mean().item()),
        "reference_ms": REFERENCE_BENCH_MS,
        "reference_hardware": "NVIDIA RTX 4070 Ti (published pytest-benchmark median)",
        "speedup": round(REFERENCE_BENCH_MS / ms, 2),
    }


#: The reference's published compute_llr medians (ms) on 2^18 symbols at
#: Es/N0 15 dB, RTX 4070 Ti, call plus stream sync (benchmarks/bench_llr.py).
REFERENCE_LLR_MS = {("qam16", "maxlog"): 0.109, ("qam16", "exact"): 0.123,
                    ("qam64", "maxlog"): 0.103, ("qam64", "exact"): 0.169,
                    ("qam256", "maxlog"): 0.106, ("qam256", "exact"): 0.251,
                    ("psk16", "maxlog"): 0.092, ("psk16", "exact"): 0.183,
                    ("qam128cross", "maxlog"): 0.291,
                    ("qam128cross", "exact"): 0.730}  # fmt: skip


def _llr_constellation(name: str):
    from commkit_amd import Constellation

    if name == "psk16":
        return Constellation.psk(16)
    if name == "qam128cross":
        return Constellation.qam(128)
    return Constellation.qam(int(name[3:]))


def reference_llr_benchmarks(dev, reps: int = 20) -> list:
    """The reference's compute_llr benchmarks: 2^18 Gray-mapped symbols after
    AWGN at 15 dB, device input; wall time of the call plus a synchronize
    (as the reference times it), median of ``reps``."""
    import torch

    from commkit_amd import compute_llr
    from commkit_amd.impairments import apply_awgn

    out = []
    gen = np.random.default_rng(0)
    for (name, method), ref_ms in REFERENCE_LLR_MS.items():
        try:
            c = _llr_constellation(name)
        except ValueError as e:  # a constellation this build does not make
            out.append({"constellation": name, "method": method, "error": str(e)})
            continue
        bits = torch.from_numpy(gen.integers(0, 2, (1 << 18) * c.bits_per_symbol,
                                             dtype=np.int8)).to(dev)  # fmt: skip
        rx = apply_awgn(c.map(bits), esn0_db=15.0, sps=1, rng=1)
        nv = 10 ** (-15.0 / 10)
        compute_llr(rx, noise_var=nv, constellation=c, method=method)
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            compute_llr(rx, noise_var=nv, constellation=c, method=method)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        out.append({"constellation": name, "method": method, "ms": round(ms, 4),
                    "reference_ms": ref_ms, "speedup": round(ref_ms / ms, 2)})  # fmt: skip
    return out


# -----------------------------------------------------------------------------
# C2: the headline (one code at scale)
# -----------------------------------------------------------------------------


def run_single(args, dev) -> tuple[dict, bool]:
    import torch

    from commkit_amd import Constellation, MinSum, metrics
    from commkit_amd.coding.ldpc import _decode, _decode_symbols
    from commkit_amd.mapping.demap import llr_any
    from commkit_amd.operating import waterfall_db
    from commkit_amd.workloads import make_frames, seeded_code

    code = seeded_code(CODE)
    c = Constellation.qam(ORDER)
    esn0 = args.esn0 if args.esn0 is not None else round(waterfall_db(CODE, ORDER) + MARGIN_DB, 3)
    F = args.frames
    fr = make_frames(code, c, esn0, F, seed=0, device=dev)
    nv = float(fr.noise_var[0])
    rx = fr.rx.contiguous()
    method = MinSum(max_iterations=50)
    state = {}
    fused = args.tail == "fused"
    sym = rx.reshape(F, -1)

    def demap():
        state["llr"] = llr_any(c, rx, nv, "maxlog")

    def dec():
        state["two"] = _decode(code, state["llr"].reshape(F, code.n), method)

    def tail():
        state["fused"] = _decode_symbols(code, sym, c, nv, method)

    # Both tails on the same symbols: one launch (the demapping in the
    # decoder's frame load) and two (demapping kernel, then the decoder);
    # the headline is args.tail's, the other is reported beside it.
    runs = {}
    modes = ("fused", "two") if fused else ("two", "fused")
    for mode in modes[:1] if args.one_tail else modes:
        if mode == "fused":
            el, (dms,), sms = timed_steps(args.steps, args.warmup, [tail])
            runs[mode] = (el, 0.0, dms, sms)
        else:
            el, (mms, dms), sms = timed_steps(args.steps, args.warmup, [demap, dec])
            runs[mode] = (el, mms, dms, sms)
    head = "fused" if fused else "two"
    elapsed, demap_ms, decode_ms, step_ms = runs[head]
    if args.one_tail and fused:
        demap()  # the LLRs the fused load computed, for the parity check (clock stopped)
        dec()
    state["out"] = state[head]
    tails_equal = all(torch.equal(x, y) for x, y in zip(state["fused"], state["two"])) \
        if "fused" in state and "two" in state else True
    bits, app, valid, iters = state["out"]
    iters_h = iters.cpu().numpy()
    # Scored by the library's own kernels, after the clock stopped.
    ber = metrics.ber(bits.reshape(-1), fr.info.reshape(-1))
    fer = metrics.fer(bits.reshape(-1), fr.info.reshape(-1), info_bits=code.k)
    value = args.steps * F * code.k / elapsed / 1e9
    base_bytes = decoder_bytes(code, iters_h)
    demap_bytes = rx.numel() * (8 + 4 * c.bits_per_symbol)
    fused_bytes = base_bytes + rx.numel() * 8 - F * code.n * 4  # symbols in instead of LLRs
    measured = F == FRAMES and args.esn0 is None

    def decoder_roofline(mode: str, ms: float) -> dict:
        prof = PROFILED_TRAFFIC["c2_fused" if mode == "fused" else "c2"]
        return roofline(fused_bytes if mode == "fused" else base_bytes, ms,
                        "ldpc_qc_layered (demap fused into the load + decode)" if mode == "fused"
                        else "ldpc_qc_layered (decode)",
                        traffic=prof["bytes"] if measured else None, source=prof["profile"],
                        limiter=LIMITER["c2"])  # fmt: skip

    def demapper_roofline(ms: float) -> dict:
        prof = PROFILED_TRAFFIC["c2"]
        return roofline(demap_bytes, ms, "llr_separable_kernel (demap)",
                        traffic=prof["demap_bytes"] if measured else None,
                        source=prof["profile"], limiter=LIMITER["c2_demap"])  # fmt: skip

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gb/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded code, uniform info bits, AWGN)",
        "config": {
            "workload": "rate-4/5 QC-LDPC n=18000 Z=360 (seed 104), 4096 codewords, Gray "
                        "64-QAM, max-log LLRs, layered NMS 0.875, <=50 iterations",
            "frames": F,
            "esn0_db": esn0,
            "code": CODE,
            "constellation": f"{ORDER}-QAM",
            "tail": ("one launch: max-log demapping fused into the decoder's frame load"
                     if fused else "two launches: demapping kernel, then the decoder"),
        },
        "frame_latency_ms": pct(step_ms),
        "post_fec_ber": ber,
        "fer": fer,
        "mean_iterations": float(iters_h.mean()),
        "kernels_ms": ({"demap+decode (fused)": round(decode_ms, 4)} if fused else
                       {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4)}),
        "roofline": decoder_roofline(head, decode_ms),
    }
    if not fused:
        line["demap_roofline"] = demapper_roofline(demap_ms)
    line["tail_roofline"] = {
        "achieved": round((fused_bytes if fused else base_bytes + demap_bytes)
                          / ((demap_ms + decode_ms) * 1e-3) / 1e9, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
    }
    other = "two" if fused else "fused"
    if other in runs:
        o_el, o_mms, o_dms, o_sms = runs[other]
        alt = {
            "tail": ("two launches: demapping kernel, then the decoder" if other == "two"
                     else "one launch: max-log demapping fused into the decoder's frame load"),
            "value": round(args.steps * F * code.k / o_el / 1e9, 3),
            "ms_per_step": round(o_el / args.steps * 1e3, 3),
            "frame_latency_ms": pct(o_sms),
            "kernels_ms": ({"demap": round(o_mms, 4), "decode": round(o_dms, 4)}
                           if other == "two" else {"demap+decode (fused)": round(o_dms, 4)}),
            "roofline": decoder_roofline(other, o_dms),
            "outputs_equal": tails_equal,
        }
        if other == "two":
            alt["demap_roofline"] = demapper_roofline(o_mms)
        line["other_tail"] = alt
    line["hbm_copy_GBs"] = round(copy_bandwidth(dev), 1)
    progress(leg="c2", value=line["value"], decode_ms=round(decode_ms, 4))
    # Parity: the oracle decodes every frame's GPU LLRs (for the fused tail,
    # those of the demapping kernel, which the fused load reproduces bit for
    # bit); the GPU LLRs of a sample of symbols against the FP64 definition.
    llr_h = state["llr"].reshape(F, code.n).cpu().numpy()
    t0 = time.perf_counter()
    want = oracle_decode(code, llr_h, method)
    parity = compare(to_host(state["out"]), want, exact=True)
    parity["oracle_s"] = round(time.perf_counter() - t0, 2)
    line["parity"] = parity
    sample = rx.reshape(-1)[: 1 << 16].cpu().numpy()
    line["llr_parity"] = llr_parity(c, sample, nv, "maxlog",
                                    state["llr"].reshape(-1)[: sample.size * c.bits_per_symbol]
                                    .cpu().numpy())  # fmt: skip
    progress(leg="c2", parity=parity["identical"], frames=parity["frames"])
    if not args.no_cpu_baseline:
        tables = oracle_tables(c)
        rx_h = fr.rx.cpu().numpy()
        probe = min(16, F)

        def run(frames: int):
            def go(reps: int):
                for _ in range(reps):
                    x = rx_h[:frames].reshape(-1)
                    llr = oracle_demap(c, x, nv, "maxlog", tables).reshape(frames, code.n)
                    oracle_decode(code, llr, method)
            return go

        t0 = time.perf_counter()
        run(probe)(1)
        per_frame = max(time.perf_counter() - t0, 1e-3) / probe
        frames = int(min(F, max(probe, args.cpu_seconds / 5 / per_frame)))
        dt, _ = time_cpu(run(frames), args.cpu_seconds, runs=5)
        line["cpu_baseline"] = cpu_line(frames * code.k / dt / 1e9,
                                        f"{frames} of the {F} frames (max-log LLRs + layered "
                                        f"min-sum), {dt:.2f} s per run")  # fmt: skip
    return line, parity["identical"] == parity["frames"] and tails_equal


# -----------------------------------------------------------------------------
# C1: the reference's plain CPU-path configuration
# -----------------------------------------------------------------------------


def run_plain(args, dev) -> tuple[dict, bool]:
    """64 codewords of the rate-1/2 Z = 27 code (n = 648), Gray 16-QAM at
    Es/N0 8 dB, max-log LLRs, flooding normalized min-sum (scale 0.75, at
    most 20 iterations, early stop).  8 dB lies below this code's 16-QAM
    waterfall (9.2 dB), so many frames fail; the line reports the goodput
    (info bits of frames decoded correctly) beside the raw rate.  The
    headline tail demaps in the flooding kernel's frame load
    (ck_ldpc_decode_symbols); the two launches are timed beside it
    (``other_tail``) and their outputs checked equal."""
    import torch

    from commkit_amd import Constellation, MinSum, metrics
    from commkit_amd.coding.ldpc import _decode, _decode_symbols
    from commkit_amd.mapping.demap import llr_any
    from commkit_amd.workloads import make_frames, seeded_code

    name, order, esn0, F = "r1/2-z27", 16, 8.0, 64
    code = seeded_code(name)
    c = Constellation.qam(order)
    fr = make_frames(code, c, esn0, F, seed=0, device=dev)
    nv = float(fr.noise_var[0])
    rx = fr.rx.contiguous()
    method = MinSum(scale=0.75, max_iterations=20, schedule="flooding")
    state = {}

    def demap():
        state["llr"] = llr_any(c, rx, nv, "maxlog")

    def dec():
        state["out"] = _decode(code, state["llr"].reshape(F, code.n), method)

    sym = rx.reshape(F, -1)

    def fused():  # the headline tail: demapping in the flooding kernel's load
        state["fused"] = _decode_symbols(code, sym, c, nv, method)

    steps = max(args.steps, 100)
    elapsed_f, (fused_ms,), step_ms_f = timed_steps(steps, max(args.warmup, 5), [fused])
    elapsed, (demap_ms, decode_ms), step_ms = timed_steps(steps, max(args.warmup, 5),
                                                          [demap, dec])  # fmt: skip
    tails_equal = all(torch.equal(a, b) for a, b in zip(state["fused"], state["out"]))
    got = to_host(state["out"])
    bits, _, _, iters = got
    ber = metrics.ber(state["out"][0].reshape(-1), fr.info.reshape(-1))
    ok = (bits == fr.info.cpu().numpy()).all(axis=1)
    per_step = elapsed_f / steps
    per_step_two = elapsed / steps
    dec_bytes = decoder_bytes(code, iters)
    fused_bytes = dec_bytes - F * (4 * code.n - 8 * code.n // c.bits_per_symbol)
    out = {
        "workload": f"{F} codewords of {name} (n={code.n}, k={code.k}), Gray {order}-QAM at "
                    f"Es/N0 {esn0} dB, max-log LLRs, flooding NMS 0.75, <=20 iterations",
        "value": round(F * code.k / per_step / 1e9, 4),
        "unit": "Gb/s",
        "goodput_gbps": round(int(ok.sum()) * code.k / per_step / 1e9, 4),
        "steps": steps,
        "ms_per_step": round(per_step * 1e3, 4),
        "frame_latency_ms": pct(step_ms_f),
        "post_fec_ber": ber,
        "fer": float(1.0 - ok.mean()),
        "mean_iterations": float(iters.mean()),
        "tail": "fused: ck_ldpc_decode_symbols, the flooding kernel's frame load demaps",
        "kernels_ms": {"demap+decode (fused)": round(fused_ms, 4)},
        "roofline": roofline(fused_bytes, fused_ms,
                             "ldpc_qc_flooding (demap fused into the load + decode)",
                             limiter="launch latency: 64 frames fill a quarter of the chip's "
                                     "CUs for a few microseconds"),  # fmt: skip
        "other_tail": {
            "tail": "two launches: demapping kernel, then the decoder",
            "value": round(F * code.k / per_step_two / 1e9, 4),
            "ms_per_step": round(per_step_two * 1e3, 4),
            "frame_latency_ms": pct(step_ms),
            "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4)},
            "roofline": roofline(dec_bytes, decode_ms, "ldpc_qc_flooding (decode)"),
            "outputs_equal": bool(tails_equal),
        },
    }
    want = oracle_decode(code, state["llr"].reshape(F, code.n).cpu().numpy(), method)
    out["parity"] = compare(got, want, exact=True)
    out["llr_parity"] = llr_parity(c, rx.reshape(-1).cpu().numpy(), nv, "maxlog",
                                   state["llr"].cpu().numpy())  # fmt: skip
    if not args.no_cpu_baseline:
        tables = oracle_tables(c)
        rx_h = rx.reshape(-1).cpu().numpy()

        def go(reps: int):
            for _ in range(reps):
                llr = oracle_demap(c, rx_h, nv, "maxlog", tables).reshape(F, code.n)
                oracle_decode(code, llr, method)

        dt, reps = time_cpu(go, args.cpu_seconds_sub)
        out["cpu_baseline"] = cpu_line(F * code.k / dt / 1e9,
                                       f"all {F} frames (max-log LLRs + flooding min-sum) "
                                       f"{reps} times a run")  # fmt: skip
    progress(leg="c1", value=out["value"], parity=out["parity"]["identical"])
    return out, out["parity"]["identical"] == F and tails_equal


# -----------------------------------------------------------------------------
# C3: mixed codes and modulations in one persistent batched launch
# -----------------------------------------------------------------------------


def run_mixed(args, dev, frames: int = 10_000) -> tuple[dict, bool]:
    """Per-frame-noise demapping of 12 (code, QAM) groups into the code
    classes (one captured graph), then one persistent batched decode (a
    captured HIP graph).  Also times the same frames decoded per code."""
    import torch

    from commkit_amd import Constellation, MinSum
    from commkit_amd.batch import BatchDecoder
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import llr_rows
    from commkit_amd.operating import waterfall_db
    from commkit_amd.workloads import MIXED_CODES, MIXED_ORDERS, make_mixed

    mx = make_mixed(frames, seed=0, device=dev)
    method = MinSum(max_iterations=50)
    bd = BatchDecoder(mx.codes, mx.rows, method, device=dev, graph=True)
    consts = {o: Constellation.qam(o) for o in MIXED_ORDERS}
    inv = [torch.from_numpy(1.0 / np.maximum(g.frames.noise_var, 1e-20)).to(dev)
           for g in mx.groups]  # fmt: skip

    def demap():
        for g, iv in zip(mx.groups, inv):
            llr_rows(consts[g.order], g.frames.rx, None, "maxlog",
                     out=bd.inputs[g.code_index][g.row0:g.row1], inv_s2=iv)  # fmt: skip

    for _ in range(max(args.warmup, 1)):
        demap()
        bd.run()
    torch.cuda.synchronize()
    eager = [x.clone() for x in bd.inputs]
    demap_graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(demap_graph):
        demap()
    torch.cuda.synchronize()
    elapsed, (demap_ms, decode_ms), step_ms = timed_steps(args.steps, 0,
                                                          [demap_graph.replay, bd.run])  # fmt: skip
    demap_same = all(torch.equal(x, y) for x, y in zip(bd.inputs, eager))
    stream = torch.cuda.current_stream(dev)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per_code = [_decode(c, bd.inputs[i], method) for i, c in enumerate(mx.codes)]
    a.record(stream)
    per_code = [_decode(c, bd.inputs[i], method) for i, c in enumerate(mx.codes)]
    b.record(stream)
    b.synchronize()
    per_code_ms = a.elapsed_time(b)
    same = all(torch.equal(bd.outputs[i].bits, per_code[i][0]) and
               torch.equal(bd.outputs[i].iterations, per_code[i][3]) for i in range(len(mx.codes)))

    errors = failed = 0
    good_bits = 0
    above = {"frames": 0, "failed": 0, "errors": 0, "bits": 0, "good_bits": 0}
    for g in mx.groups:
        code = mx.codes[g.code_index]
        bits = bd.outputs[g.code_index].bits[g.row0:g.row1]
        e = (bits != g.frames.info).sum(dim=1).cpu().numpy()
        errors += int(e.sum())
        failed += int((e > 0).sum())
        good_bits += int((e == 0).sum()) * code.k
        hi = g.frames.esn0_db >= waterfall_db(MIXED_CODES[g.code_index], g.order)
        above["frames"] += int(hi.sum())
        above["failed"] += int((e[hi] > 0).sum())
        above["errors"] += int(e[hi].sum())
        above["bits"] += int(hi.sum()) * code.k
        above["good_bits"] += int((e[hi] == 0).sum()) * code.k
    dec_bytes = sum(decoder_bytes(c, bd.outputs[i].iterations.cpu().numpy())
                    for i, c in enumerate(mx.codes))  # fmt: skip
    demap_bytes = sum(g.frames.rx.numel() * (8 + 4 * int(np.log2(g.order))) for g in mx.groups)
    per_step = elapsed / args.steps
    out = {
        "workload": f"mixed: {frames} frames, codes {'/'.join(MIXED_CODES)} x Gray "
                    "4/16/64/256-QAM uniform, Es/N0 ~ U[wf-0.5, wf+1] dB per frame, max-log LLRs, "
                    "layered NMS 0.875, <=50 iterations, one batched launch (HIP graph)",
        "value": round(mx.info_bits / per_step / 1e9, 3),
        "unit": "Gb/s",
        "goodput_gbps": round(good_bits / per_step / 1e9, 3),
        "frames": frames,
        "rows_per_code": mx.rows,
        "steps": args.steps,
        "ms_per_step": round(per_step * 1e3, 3),
        "frame_latency_ms": pct(step_ms),
        "post_fec_ber": errors / mx.info_bits,
        "fer": failed / frames,
        "at_or_above_waterfall": {
            "frames": above["frames"],
            "fer": above["failed"] / max(above["frames"], 1),
            "post_fec_ber": above["errors"] / max(above["bits"], 1),
            "goodput_gbps": round(above["good_bits"] / per_step / 1e9, 3),
        },
        "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4),
                       "decode_per_code_launches": round(per_code_ms, 4)},
        "batch_equals_per_code": bool(same),
        "demap_graph_equals_eager": bool(demap_same),
        "roofline": roofline(dec_bytes, decode_ms, "ldpc_batch (decode)"),
        "demap_roofline": roofline(demap_bytes, demap_ms, "llr_separable_rows x 12 (demap)"),
    }
    progress(leg="c3", value=out["value"], decode_ms=round(decode_ms, 4))
    # Parity on every frame: each (code, QAM) group decoded again by the C
    # oracle from the LLRs the demap graph wrote (about 10 s on 16 threads).
    ok_all = True
    sample_frames = 0
    identical = 0
    divergent = []
    for g in mx.groups:
        code = mx.codes[g.code_index]
        r0, r1 = g.row0, g.row1
        o = bd.outputs[g.code_index]
        got = (o.bits[r0:r1].cpu().numpy(), o.llrs[r0:r1].cpu().numpy(),
               o.valid[r0:r1].cpu().numpy().astype(bool), o.iterations[r0:r1].cpu().numpy())
        want = oracle_decode(code, bd.inputs[g.code_index][r0:r1].cpu().numpy(), method)
        p = compare(got, want, np.arange(r0, r1), exact=True)
        sample_frames += p["frames"]
        identical += p["identical"]
        divergent += [f"{MIXED_CODES[g.code_index]}:{f}" for f in p["divergent"]]
    out["parity"] = {"frames": sample_frames, "identical": identical, "divergent": divergent[:32],
                     "rule": "bits, APP, valid and iterations equal",
                     "sample": "every frame of the 12 (code, QAM) groups"}  # fmt: skip
    ok_all = identical == sample_frames
    if not args.no_cpu_baseline:
        tables = {o: oracle_tables(consts[o]) for o in MIXED_ORDERS}
        picks = []
        for g in mx.groups:
            s = min(8, g.row1 - g.row0)
            rx_h = g.frames.rx[:s].cpu().numpy()
            picks.append((g, rx_h, g.frames.noise_var[:s]))
        bits_total = sum(mx.codes[g.code_index].k * rx.shape[0] for g, rx, _ in picks)

        def go(reps: int):
            for _ in range(reps):
                for g, rx_h, nvs in picks:
                    code = mx.codes[g.code_index]
                    c = consts[g.order]
                    llr = np.stack([oracle_demap(c, rx_h[i], float(nvs[i]), "maxlog",
                                                 tables[g.order]) for i in range(rx_h.shape[0])])
                    oracle_decode(code, llr.reshape(-1, code.n), method)

        dt, reps = time_cpu(go, args.cpu_seconds_sub)
        out["cpu_baseline"] = cpu_line(bits_total / dt / 1e9,
                                       f"the first 8 frames of each of the 12 (code, QAM) groups "
                                       f"({sum(r.shape[0] for _, r, _ in picks)} frames, max-log "
                                       f"LLRs at per-frame Es/N0 + layered min-sum)")  # fmt: skip
    bd.close()
    progress(leg="c3", parity=identical, frames=sample_frames)
    return out, ok_all


# -----------------------------------------------------------------------------
# C5: shaped 256-QAM, exact LLRs with the prior, sum-product
# -----------------------------------------------------------------------------


def run_shaped(args, dev) -> tuple[dict, bool]:
    """Shaped 256-QAM (Maxwell-Boltzmann nu = 0.05) by probabilistic
    amplitude shaping over the rate-5/6 code: 2 x 2^22 received symbols
    (3,106 whole codewords of 2,700 symbols and 2,408 filler symbols)
    demapped with exact LLRs and the prior straight into the codewords' bit
    layout, layered sum-product of at most 30 iterations.  ``value`` is that
    tail's rate in the code's information bits; the CCDM dematcher that
    turns decoded amplitude bits back into payload is timed after it
    (``pas``), and the frames are served once more through the serving loop
    (``served``)."""
    import torch

    from commkit_amd import Constellation, SumProduct
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import llr_permuted
    from commkit_amd.operating import SHAPED_WATERFALL_DB
    from commkit_amd.workloads import make_pas_frames, seeded_code

    name, nu = "r5/6-z360", 0.05
    code = seeded_code(name)
    c = Constellation.qam(256).shaped(nu=nu)
    esn0 = round(SHAPED_WATERFALL_DB[(name, 256, nu)] + 0.5, 3)
    total = 2 * (1 << 22)
    S = code.n // c.bits_per_symbol
    F = total // S
    fr = make_pas_frames(code, c, esn0, F, seed=0, device=dev)
    filler = total - F * S
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    pts = torch.from_numpy(c.points.astype(np.complex64)).to(dev)
    fill = pts[torch.randint(0, c.order, (filler,), generator=g, device=dev)]
    nv = float(fr.noise_var[0])
    fill = fill + np.sqrt(nv / 2) * torch.complex(
        torch.randn(filler, generator=g, device=dev), torch.randn(filler, generator=g, device=dev))
    rx = torch.cat([fr.rx.reshape(-1), fill.to(torch.complex64)])
    method = SumProduct(max_iterations=30)
    stream = torch.cuda.current_stream(dev)
    ccdm = fr.pas.matcher
    perm32 = fr.perm.to(torch.int32).contiguous()
    state = {}

    def demap():
        state["llr"] = llr_permuted(c, rx, nv, "exact", perm32, F)

    def dec():
        state["out"] = _decode(code, state["llr"], method)

    def dematch(fec_bits):
        payload, ok = ccdm._dematch_labels(fec_bits[:, : ccdm.n])
        return torch.cat([payload, fec_bits[:, ccdm.n:]], dim=1), ok

    elapsed, (demap_ms, decode_ms), step_ms = timed_steps(args.steps, args.warmup, [demap, dec])
    bits, _, valid, iters = state["out"]
    dematch(bits)
    dm_ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    dm_ev[0].record(stream)
    payload, dm_ok = dematch(bits)
    dm_ev[1].record(stream)
    dm_ev[1].synchronize()
    dematch_ms = dm_ev[0].elapsed_time(dm_ev[1])
    errors = (payload != fr.info).sum(dim=1).cpu().numpy()
    fec_errors = int((bits != fr.sent[:, fr.iperm][:, : code.k]).sum().item())
    flagged = int((~(valid & dm_ok)).sum().item())
    iters_h = iters.cpu().numpy()
    dec_bytes = decoder_bytes(code, iters_h)
    demap_bytes = total * (8 + 4 * c.bits_per_symbol)
    per_step = elapsed / args.steps
    prof = PROFILED_TRAFFIC["c5"]
    out = {
        "workload": f"shaped: 2 x 2^22 symbols of 256-QAM shaped(nu={nu}), {F} PAS codewords of "
                    f"{name} + {filler} filler symbols, Es/N0 {esn0} dB (shaped waterfall + 0.5), "
                    "exact LLRs with prior, layered sum-product <=30 iterations",
        "value": round(F * code.k / per_step / 1e9, 3),
        "unit": "Gb/s",
        "frames": F,
        "esn0_db": esn0,
        "steps": args.steps,
        "ms_per_step": round(per_step * 1e3, 3),
        "frame_latency_ms": pct(step_ms),
        "post_fec_ber": fec_errors / (F * code.k),
        "fer": float((errors > 0).mean()),
        "mean_iterations": float(iters_h.mean()),
        "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4)},
        "pas": {
            "payload_bits_per_codeword": fr.pas.k,
            "dematch_ms": round(dematch_ms, 4),
            "payload_ber": float(errors.sum() / (F * fr.pas.k)),
            "frames_flagged": flagged,
            "payload_gbps_with_dematch": round(F * fr.pas.k / (per_step + dematch_ms * 1e-3) / 1e9,
                                               3),
        },
        "roofline": roofline(dec_bytes, decode_ms, "ldpc_qc_layered<sum-product> (decode)",
                             traffic=prof["bytes"], source=prof["profile"],
                             limiter=LIMITER["c5"]),  # fmt: skip
        "demap_roofline": roofline(demap_bytes, demap_ms,
                                   "llr_separable_kernel_w4<4,false> (exact demap, permuted store)",
                                   traffic=prof["demap_bytes"], source=prof["profile"],
                                   limiter=LIMITER["c5_demap"]),
    }
    progress(leg="c5", value=out["value"], decode_ms=round(decode_ms, 4))
    # Parity: the oracle decodes the GPU LLRs of every codeword (sum-product:
    # identical hard decisions and iterations, APP deviation reported); the
    # exact LLRs of a sample of symbols against FP64.
    s = F
    got = tuple(x[:s] for x in to_host(state["out"]))
    want = oracle_decode(code, state["llr"][:s].cpu().numpy(), method)
    out["parity"] = compare(got, want, exact=False)
    out["parity"]["sample"] = f"all {s} codewords"
    xs = fr.rx.reshape(-1)[: 1 << 14]
    from commkit_amd.mapping.demap import llr_any

    out["llr_parity"] = llr_parity(c, xs.cpu().numpy(), nv, "exact",
                                   llr_any(c, xs, nv, "exact").cpu().numpy(),
                                   floor=2.0 ** -5)  # fmt: skip
    progress(leg="c5", parity=out["parity"]["identical"], frames=s)
    if not args.no_cpu_baseline:
        tables = oracle_tables(c)
        iperm = fr.iperm.cpu().numpy()
        k_cw = 8
        rx_h = fr.rx[:k_cw].reshape(-1).cpu().numpy()

        def go(reps: int):
            for _ in range(reps):
                llr = oracle_demap(c, rx_h, nv, "exact", tables).reshape(k_cw, code.n)
                oracle_decode(code, np.ascontiguousarray(llr[:, iperm]), method)

        dt, reps = time_cpu(go, args.cpu_seconds_sub)
        out["cpu_baseline"] = cpu_line(k_cw * code.k / dt / 1e9,
                                       f"the first {k_cw} codewords (exact LLRs with the prior, "
                                       "gathered to codeword order, + layered sum-product)")  # fmt: skip
    out["served"] = shaped_served(args, code, c, fr, nv, F, method, dev)
    return out, True


def shaped_served(args, code, c, fr, nv: float, F: int, method, dev) -> dict:
    """The shaped frames at saturation inside the serving loop: requests of
    U{1..256} whole codewords, all queued at once; each request's symbols are
    demapped (exact, with the prior) straight into its ring slot, in codeword
    bit order, and decoded by continuous batching; decoded bits are copied
    out of the ring as requests complete, and one dematcher launch over all
    frames closes the run (its time is inside the measured wall clock)."""
    import torch

    from commkit_amd.mapping.demap import SegmentsDemapper, llr_permuted
    from commkit_amd.serve import Schedule, Server, drive

    rng = np.random.Generator(np.random.PCG64(11))
    sizes = []
    while sum(sizes) < F:
        sizes.append(int(min(rng.integers(1, 257), F - sum(sizes))))
    sizes = np.asarray(sizes)
    first = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    n_req = sizes.size
    sched = Schedule(np.zeros(n_req), np.arange(n_req) % 32, np.zeros(n_req, np.int64), sizes,
                     np.zeros(n_req, np.int64))  # fmt: skip
    S = code.n // c.bits_per_symbol
    rx = fr.rx.reshape(-1)
    perm32 = fr.perm.to(device=dev, dtype=torch.int32).contiguous()

    seg_dm = SegmentsDemapper(c, "exact", dev) if args.demap_per_turn else None

    def fill(slot, i):
        a, f = int(first[i]), int(sizes[i])
        if seg_dm is not None:
            seg_dm.add(rx[a * S:(a + f) * S], slot.llrs, noise_var=nv, perm=perm32)
        else:
            llr_permuted(c, rx[a * S:(a + f) * S], nv, "exact", perm32, f, out=slot.llrs)

    flush = seg_dm.flush if seg_dm is not None else None

    bits_out = torch.empty((F, code.k), dtype=torch.int8, device=dev)
    iters_out = torch.empty(F, dtype=torch.int32, device=dev)

    def on_done(slot, i):
        a, f = int(first[i]), int(sizes[i])
        bits_out[a:a + f].copy_(slot.bits)
        iters_out[a:a + f].copy_(slot.iterations)

    # Launched in two halves by default, so the first half decodes while the
    # rest is admitted (one launch of every frame: 14.2 Gb/s, p50 3.8 ms;
    # halves: 15.0 Gb/s, p50 2.4 ms).
    cap = args.max_launch_frames or (F + 1) // 2
    srv = Server([code], method, hbm_budget=16 << 30, max_inflight=args.max_inflight, device=dev,
                 max_launch_frames=cap)  # fmt: skip
    warm = Schedule(np.zeros(4), np.arange(4), np.zeros(4, np.int64), sizes[:4], np.zeros(4, np.int64))
    drive(srv, warm, fill, before_step=flush)
    t0 = time.perf_counter()
    res = drive(srv, sched, fill, on_done=on_done, before_step=flush)
    ccdm = fr.pas.matcher
    payload, dm_ok = ccdm._dematch_labels(bits_out[:, : ccdm.n])
    payload = torch.cat([payload, bits_out[:, ccdm.n:]], dim=1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    srv.close()
    errors = int((payload != fr.info).sum().item())
    it_sum = int(iters_out.sum(dtype=torch.int64).item())
    E = code.checks.size
    dec_bytes = F * (8 * code.n + code.k + 5) + 8.0 * E * it_sum
    tail_bytes = dec_bytes + F * S * (8 + 4 * c.bits_per_symbol)
    progress(leg="c5 served", gbps=round(res.gbps, 3))
    return {
        "gbps": round(res.gbps, 3),
        "payload_gbps_with_dematch": round(F * fr.pas.k / elapsed / 1e9, 3),
        "requests": int(n_req),
        "frames": int(F),
        "launches": res.launches,
        "max_launch_frames": cap,
        "max_inflight": args.max_inflight,
        "frame_latency_ms": {"p50": round(res.percentile(50), 3),
                             "p99": round(res.percentile(99), 3)},
        "payload_ber": errors / (F * fr.pas.k),
        "frames_flagged": int((~dm_ok).sum().item()),
        "mean_iterations": it_sum / F,
        "tail_hbm_frac": round(tail_bytes / res.elapsed_s / 1e9 / HBM_PEAK_GBS, 4),
        "note": "gbps: the code's information bits over the wall clock of the loop (host "
                "admission, each request demapped into its slot, batched decode); the payload "
                "rate adds the dematcher launch after the loop",
    }


# -----------------------------------------------------------------------------
# C4: the serving loop
# -----------------------------------------------------------------------------


def run_serving(args, dev) -> tuple[dict, bool]:
    """32 Poisson request streams (stream s: code s mod 3, Gray QAM order
    4/16/64/256 by s // 3, request sizes uniform in 1..256 frames, Es/N0 a
    reflected random walk of 0.1 dB steps over [wf - 0.5, wf + 1]) served by
    the continuous-batching server under a 16 GiB HBM budget.  A request
    carries received symbols (from pools of pre-generated frames per (code,
    QAM, Es/N0 step)); admission demaps them (max-log, the request's noise
    variance) straight into the request's ring slot, and the server decodes
    the slots by continuous batching.  First the saturated throughput (every
    request queued at once), then each offered load as a fraction of it;
    per-frame latency is completion seen by the host minus arrival."""
    import torch

    from commkit_amd import Constellation, MinSum
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import RowsDemapper, SegmentsDemapper, llr_rows
    from commkit_amd.operating import waterfall_db
    from commkit_amd.serve import Schedule, Server, burst_schedule, drive, poisson_schedule
    from commkit_amd.workloads import MIXED_CODES, MIXED_ORDERS, make_frames, seeded_code

    streams, bins, max_frames = 32, 16, 256
    codes = [seeded_code(n) for n in MIXED_CODES]
    pairs = [(s % 3, MIXED_ORDERS[(s // 3) % 4]) for s in range(streams)]
    consts = {o: Constellation.qam(o) for o in MIXED_ORDERS}
    pools: dict = {}
    for ci, order in sorted(set(pairs)):
        c = consts[order]
        wf = waterfall_db(MIXED_CODES[ci], order)
        per_bin = []
        for b in range(bins):
            fr = make_frames(codes[ci], c, wf - 0.5 + 0.1 * b, max_frames,
                             seed=100 * ci + order + 7 * b, device=dev)  # fmt: skip
            inv = torch.full((max_frames,), 1.0 / float(fr.noise_var[0]), dtype=torch.float64,
                             device=dev)  # fmt: skip
            per_bin.append((fr.rx.contiguous(), inv, float(fr.noise_var[0])))
        pools[(ci, order)] = per_bin
    torch.cuda.synchronize()
    progress(leg="c4", pools=len(pools))
    method = MinSum(max_iterations=50)
    srv = Server(codes, method, hbm_budget=16 << 30, max_inflight=args.max_inflight, device=dev,
                 max_launch_frames=args.max_launch_frames or None)  # fmt: skip
    mean_bits = float(np.mean([codes[ci].k for ci, _ in pairs])) * (1 + max_frames) / 2

    # Each admitted request is demapped into its ring slot by a launch of
    # its own, or with --demap-per-turn queued per constellation and
    # launched once per loop turn (one launch per 16 requests).
    if args.demap_per_turn:
        rows_dm = {}
        seg_dm = {o: SegmentsDemapper(consts[o], "maxlog", dev) for o in MIXED_ORDERS}
    else:
        rows_dm = {o: RowsDemapper(consts[o], "maxlog", dev) for o in MIXED_ORDERS}
        seg_dm = {}

    def flush_all():
        for d in seg_dm.values():
            d.flush()

    # Device addresses of each pool's symbols and noise, for the fills.
    pool_ptrs = {key: [(rx.data_ptr(), inv.data_ptr(), rx.shape[1]) for rx, inv, _ in per_bin]
                 for key, per_bin in pools.items()}  # fmt: skip

    def make_fill(sched: Schedule):
        stream_pairs = [pairs[int(s)] for s in sched.stream]
        snr_bins = sched.snr_bin.tolist()

        def fill(slot, i):
            ci, order = stream_pairs[i]
            f = slot.frames
            if seg_dm:
                rx, inv, _ = pools[(ci, order)][snr_bins[i]]
                seg_dm[order].add(rx[:f], slot.llrs, inv_s2=inv[:f])
            else:
                x_ptr, inv_ptr, S = pool_ptrs[(ci, order)][snr_bins[i]]
                rows_dm[order].raw(x_ptr, f, S, inv_ptr, slot.llrs_ptr)
        return fill

    def codes_of(sched: Schedule) -> Schedule:
        sched.code = np.asarray([pairs[int(s)][0] for s in sched.stream])
        return sched

    # Saturated: a burst of requests all present at the start.  The first
    # requests' outputs are kept for the parity check.
    n_sat = 3000
    warm = codes_of(burst_schedule(64, streams=streams, num_codes=3, max_frames=max_frames,
                                   snr_bins=bins, seed=2))  # fmt: skip
    drive(srv, warm, make_fill(warm), before_step=flush_all)
    burst = codes_of(burst_schedule(n_sat, streams=streams, num_codes=3, max_frames=max_frames,
                                    snr_bins=bins, seed=1))  # fmt: skip
    kept: dict = {}

    def keep(slot, i):
        if i < 12:
            kept[i] = (slot.llrs.clone(), slot.bits.clone(), slot.app.clone(),
                       slot.valid.clone(), slot.iterations.clone())  # fmt: skip

    sat = drive(srv, burst, make_fill(burst), on_done=keep, before_step=flush_all)
    # Algorithmic bytes of the saturated run: a frame's iteration count is
    # the same alone and in any batch, so pool frames are decoded per call,
    # outside the timed region, for the counts.
    iters_cum = {}
    for key, per_bin in pools.items():
        ci, order = key
        cum = []
        for rx, inv, _ in per_bin:
            llr = llr_rows(consts[order], rx, None, "maxlog", inv_s2=inv)
            cum.append(np.cumsum(_decode(codes[ci], llr, method)[3].cpu().numpy()))
        iters_cum[key] = cum
    sat_bytes = 0.0
    demap_bytes = 0.0
    for i in range(len(burst)):
        ci, order = pairs[int(burst.stream[i])]
        f = int(burst.frames[i])
        cd = codes[ci]
        its = int(iters_cum[(ci, order)][int(burst.snr_bin[i])][f - 1])
        sat_bytes += f * (8 * cd.n + cd.k + 5) + 8.0 * cd.checks.size * its
        demap_bytes += f * (cd.n // int(np.log2(order))) * (8 + 4 * int(np.log2(order)))
    progress(leg="c4", saturated_gbps=round(sat.gbps, 3), launches=sat.launches)
    loads = []
    fracs = [float(x) for x in args.loads.split(",") if x]
    if args.bound_load not in fracs:
        fracs = sorted(fracs + [args.bound_load])
    for frac in fracs:
        rate = frac * sat.gbps * 1e9 / mean_bits
        sched = codes_of(poisson_schedule(rate, args.serve_seconds, streams=streams, num_codes=3,
                                          max_frames=max_frames, seed=int(frac * 1000)))  # fmt: skip
        res = drive(srv, sched, make_fill(sched), before_step=flush_all)
        p99 = round(res.percentile(99), 3)
        loads.append({
            "offered_fraction": frac,
            "offered_gbps": round(frac * sat.gbps, 3),
            "requests": res.requests,
            "frames": res.frames,
            "carried_gbps": round(res.gbps, 3),
            "frame_latency_ms": {"p50": round(res.percentile(50), 3), "p99": p99},
            "p99_bound_ms": args.p99_bound_ms,
            "met": p99 <= args.p99_bound_ms,
            "admissions_deferred": res.rejected,
            "launches": res.launches,
        })
        progress(leg="c4", load=loads[-1])
    srv.close()
    at = next(x for x in loads if x["offered_fraction"] == args.bound_load)
    bound = {"offered_fraction": args.bound_load, "offered_gbps": at["offered_gbps"],
             "p99_bound_ms": args.p99_bound_ms, "p99_ms": at["frame_latency_ms"]["p99"],
             "met": at["frame_latency_ms"]["p99"] <= args.p99_bound_ms}
    out = {
        "workload": "serving: 32 Poisson streams, codes r5/6-z360/r2/3-z96/r1/2-z27, Gray "
                    "4/16/64/256-QAM, requests U{1..256} frames of received symbols demapped "
                    "into the ring on admission, Es/N0 random walk in [wf-0.5, wf+1] dB, 16 GiB "
                    "HBM budget, layered NMS 0.875 <=50 iterations, continuous batching",
        "value": round(sat.gbps, 3),
        "unit": "Gb/s",
        "max_inflight": args.max_inflight,
        "max_launch_frames": srv.max_launch_frames,
        "serve_seconds": args.serve_seconds,
        "latency_bound": bound,
        "saturated": {"gbps": round(sat.gbps, 3), "requests": sat.requests,
                      "frames": sat.frames, "launches": sat.launches,
                      "tail_hbm_frac": round((sat_bytes + demap_bytes) / sat.elapsed_s / 1e9 /
                                             HBM_PEAK_GBS, 4),
                      "frame_latency_ms": {"p50": round(sat.percentile(50), 3),
                                           "p99": round(sat.percentile(99), 3)}},
        "loads": loads,
        "roofline": {"bound": "hbm", "achieved": round((sat_bytes + demap_bytes) / sat.elapsed_s
                                                       / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round((sat_bytes + demap_bytes) / sat.elapsed_s / 1e9 /
                                   HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "llr_separable_rows + ldpc_batch over the "
                                                 "saturated loop's wall clock"},
    }
    # Parity: the first served requests' demapped LLRs decoded by the oracle.
    frames = identical = 0
    divergent = []
    for i, (llr, bits, app, valid, its) in sorted(kept.items()):
        ci, _ = pairs[int(burst.stream[i])]
        want = oracle_decode(codes[ci], llr.cpu().numpy(), method)
        got = (bits.cpu().numpy(), app.cpu().numpy(), valid.cpu().numpy().astype(bool),
               its.cpu().numpy())
        p = compare(got, want, exact=True)
        frames += p["frames"]
        identical += p["identical"]
        divergent += [f"request {i}: {f}" for f in p["divergent"]]
    out["parity"] = {"frames": frames, "identical": identical, "divergent": divergent[:32],
                     "rule": "bits, APP, valid and iterations equal",
                     "sample": f"the {len(kept)} first requests of the saturated burst, "
                               "decoded by the oracle from the LLRs the loop wrote"}  # fmt: skip
    if not args.no_cpu_baseline:
        picks = []
        for (ci, order), per_bin in sorted(pools.items()):
            rx, _, nv = per_bin[bins // 2]
            picks.append((ci, order, rx[:4].cpu().numpy(), nv))
        tables = {o: oracle_tables(consts[o]) for o in MIXED_ORDERS}
        bits_total = sum(codes[ci].k * rx.shape[0] for ci, _, rx, _ in picks)

        def go(reps: int):
            for _ in range(reps):
                for ci, order, rx_h, nv in picks:
                    llr = oracle_demap(consts[order], rx_h.reshape(-1), nv, "maxlog",
                                       tables[order]).reshape(-1, codes[ci].n)  # fmt: skip
                    oracle_decode(codes[ci], llr, method)

        dt, _ = time_cpu(go, args.cpu_seconds_sub)
        out["cpu_baseline"] = cpu_line(bits_total / dt / 1e9,
                                       "4 frames of each of the 12 (code, QAM) pools at their "
                                       "middle Es/N0 step (max-log LLRs + layered min-sum)")  # fmt: skip
    return out, identical == frames


# -----------------------------------------------------------------------------


def main() -> int:
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(frames shard by whole codewords across GPUs; no multi-GPU run here).",
              file=sys.stderr)  # fmt: skip
        return 2
    import torch

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        copy_calibration(dev)
        return 0
    single = {"plain": run_plain, "mixed": run_mixed, "shaped": run_shaped,
              "serving": run_serving}  # fmt: skip
    if args.workload in single:
        out, ok = single[args.workload](args, dev)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, **out}), flush=True)
        return 0 if ok else 1
    line, ok = run_single(args, dev)
    torch.cuda.empty_cache()
    if args.workload == "all":
        legs = [x for x in args.legs.split(",") if x]
        runners = {"c1": ("c1_plain", run_plain), "c3": ("c3_mixed", run_mixed),
                   "c4": ("c4_serving", run_serving), "c5": ("c5_shaped", run_shaped)}
        configs = {}
        for leg in legs:
            key, fn = runners[leg]
            sub, sub_ok = fn(args, dev)
            configs[key] = sub
            ok = ok and sub_ok
            torch.cuda.empty_cache()
        line["configs"] = configs
        if not args.no_reference:
            line["reference_benchmark"] = reference_benchmark(dev)
            line["reference_llr_benchmarks"] = reference_llr_benchmarks(dev)
    print(json.dumps(line), flush=True)
    if not ok:
        print("parity: a min-sum frame differs from the oracle (see the parity fields)",
              file=sys.stderr)  # fmt: skip
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
