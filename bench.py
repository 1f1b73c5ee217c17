"""Benchmark of the soft-decision receive tail on one MI355X.

Workload (one GPU): the rate-4/5 QC-LDPC code drawn from seed 104 (10 x 50
base matrix, column weights 3-4, shifts in [0, 360), Z = 360, n = 18000,
k = 14400); 4096 codewords of uniform info bits, encoded, Gray 64-QAM,
complex64, AWGN at 0.3 dB above the code's measured waterfall
(commkit_amd.operating).  A step is one pass of the hot path over the batch
already resident in HBM: max-log float32 LLRs of the 12,288,000 symbols
(``compute_llr``), then layered normalized min-sum (scale 0.875, at most 50
iterations, syndrome early stop) of the 4096 frames (``decode``).

Prints one JSON line: decoded info Gb/s over the timed steps, the frame
latency of a step (every frame of a batch completes with it), post-FEC BER
and FER of the decoded bits, the roofline of the dominant kernel (the
decoder: algorithmic bytes per launch over its average duration from HIP
events on the launch stream) and the CPU baseline (the oracle's C
restatement of the reference CPU path, timed on a bounded sample of the
same workload on this host's cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
CODE = "r4/5-z360"
ORDER = 64
FRAMES = 4096
MARGIN_DB = 0.3

# HBM bytes per decoder launch on the default workload, from FETCH_SIZE and
# WRITE_SIZE counters scaled by a plain copy of known size
# (profiles/ldpc_qc_layered_r45-z360_64qam_4096cw.md).  Counters cannot be
# read inside this process, so the figure is reported only for the workload
# it was measured on.
PROFILED_TRAFFIC = {"frames": FRAMES, "bytes": 2.320e9}
# The same for the sum-product decoder launch of the shaped workload (3106
# codewords; profiles/ldpc_sum_product_shaped256_3106cw.md).
PROFILED_TRAFFIC_SHAPED = {"frames": 3106, "bytes": 8.415e9}


def _parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--esn0", type=float, default=None,
                    help="override Es/N0 (dB); default: measured waterfall + 0.3 dB")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target CPU time of the baseline sample (5 timed runs)")
    ap.add_argument("--loads", type=str, default="0.3,0.6,0.9",
                    help="serving: offered loads as fractions of the saturated throughput")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="serving: seconds of arrivals per offered load")
    ap.add_argument("--max-inflight", type=int, default=2,
                    help="serving: launches in flight (one workspace and stream each)")
    ap.add_argument("--bound-load", type=float, default=0.6,
                    help="serving: offered load (fraction of saturated) the p99 bound is stated at")
    ap.add_argument("--p99-bound-ms", type=float, default=25.0,
                    help="serving: stated p99 frame-latency bound at --bound-load")
    ap.add_argument("--workload", choices=("single", "mixed", "shaped", "serving"),
                    default="single",
                    help="single: one code and constellation (default); mixed: 10,000 frames "
                         "of three codes and four QAM orders at per-frame Es/N0, one "
                         "persistent batched decode; shaped: 2 x 2^22 symbols of shaped "
                         "256-QAM, exact LLRs with the prior, layered sum-product; serving: "
                         "32 Poisson request streams through the continuous-batching server")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device-to-device twice (counter calibration) and exit")
    return ap.parse_args()


#: The reference's own published decoder benchmark (context, not the
#: metric's config): DVB-S2 rate 1/2 normal frames, 64 codewords, QPSK at
#: Es/N0 1.5 dB, layered MinSum() (commkit benchmarks/bench_coding.py:37-87),
#: median 15.745 ms on an RTX 4070 Ti.
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
                    ("psk16", "maxlog"): 0.092, ("psk16", "exact"): 0.183}  # fmt: skip


def reference_llr_benchmarks(dev, reps: int = 20) -> list:
    """The reference's compute_llr benchmarks: 2^18 Gray-mapped symbols after
    AWGN at 15 dB, device input; wall time of the call plus a synchronize
    (as the reference times it), median of ``reps``."""
    import torch

    from commkit_amd import Constellation, compute_llr
    from commkit_amd.impairments import apply_awgn

    out = []
    gen = np.random.default_rng(0)
    for (name, method), ref_ms in REFERENCE_LLR_MS.items():
        c = Constellation.psk(16) if name == "psk16" else Constellation.qam(int(name[3:]))
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


def decoder_bytes(code, iters: np.ndarray) -> float:
    """Algorithmic HBM bytes of one decode launch: per frame 4n (LLRs in) + 4n
    (APP out) + k (bits) + 5 (iterations, valid) + I_f * 8E (R read + write)."""
    E = code.checks.size
    per = 8 * code.n + code.k + 5
    return float(iters.size * per + 8.0 * E * iters.astype(np.float64).sum())


def cpu_baseline(code, constellation, rx_host: np.ndarray, info_host: np.ndarray,
                 noise_var: float, target_s: float) -> dict:
    """The oracle's C restatement (max-log LLRs + layered min-sum) on a sample."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from commkit_amd.mapping.demap import opposite_levels
    from commkit_amd.mapping.constellation import pack_bits
    from oracle import cpu

    re, im = constellation._axes
    closed = []
    for a in (re, im):
        lo, hi = opposite_levels(a.labels)
        closed.append((a.levels, pack_bits(a.labels), lo, hi, a.bits))
    inv = float(np.float32(1.0 / noise_var))
    row_ptr, cols, _ = code._schedule

    def run(frames: int) -> float:
        t0 = time.perf_counter()
        x = rx_host[:frames].reshape(-1)
        llr = cpu.llr_separable_maxlog(x, closed, inv).reshape(frames, code.n)
        full = np.zeros((frames, code.num_variables), np.float32)
        full[:, : code.n] = llr
        cpu.ldpc_layered(full, row_ptr, cols, sum_product=False, scale=0.875,
                         max_iterations=50, early_stop=True)  # fmt: skip
        return time.perf_counter() - t0

    probe = min(16, rx_host.shape[0])
    run(probe)  # warm-up
    t = run(probe)
    per_run = target_s / 5
    frames = int(min(rx_host.shape[0], max(probe, probe * per_run / max(t, 1e-3))))
    times = [run(frames) for _ in range(5)]
    dt = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            model = next(ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {
        "value": frames * code.k / dt / 1e9,
        "unit": "Gb/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{frames} of the {rx_host.shape[0]} frames (max-log LLRs + layered "
                  f"min-sum, C/OpenMP oracle, {threads} threads of {os.cpu_count()} on "
                  f"{model}), median of 5 runs, {dt:.2f} s each",
    }


def run_mixed(args: argparse.Namespace) -> dict:
    """The mixed workload: per-frame-noise demapping of 12 (code, QAM) groups
    into the code classes, then one persistent batched decode (a captured HIP
    graph).  Also times the same frames decoded per code (three launches)."""
    import torch

    from commkit_amd import Constellation, MinSum
    from commkit_amd.batch import BatchDecoder
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import llr_rows
    from commkit_amd.workloads import MIXED_CODES, MIXED_ORDERS, make_mixed

    dev = torch.device("cuda", 0)
    frames = args.frames if args.frames != FRAMES else 10_000
    mx = make_mixed(frames, seed=0, device=dev)
    method = MinSum(max_iterations=50)
    bd = BatchDecoder(mx.codes, mx.rows, method, device=dev, graph=True)
    consts = {o: Constellation.qam(o) for o in MIXED_ORDERS}
    stream = torch.cuda.current_stream(dev)

    # Per-frame 1/sigma^2 uploaded once; the twelve (code, order) demaps are
    # then captured in one graph, like the batched decode.
    inv = [torch.from_numpy(1.0 / np.maximum(g.frames.noise_var, 1e-20)).to(dev)
           for g in mx.groups]  # fmt: skip

    def demap():
        for g, iv in zip(mx.groups, inv):
            llr_rows(consts[g.order], g.frames.rx, None, "maxlog",
                     out=bd.inputs[g.code_index][g.row0:g.row1], inv_s2=iv)  # fmt: skip

    for _ in range(args.warmup):
        demap()
        bd.run()
    torch.cuda.synchronize()
    eager = [x.clone() for x in bd.inputs]
    demap_graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(demap_graph):
        demap()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    step_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        ev[s][0].record(stream)
        demap_graph.replay()
        ev[s][1].record(stream)
        bd.run()
        ev[s][2].record(stream)
        ev[s][2].synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    demap_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    decode_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))

    # The same frames decoded per code, one launch each.
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per_code = [_decode(c, bd.inputs[i], method) for i, c in enumerate(mx.codes)]
    a.record(stream)
    per_code = [_decode(c, bd.inputs[i], method) for i, c in enumerate(mx.codes)]
    b.record(stream)
    b.synchronize()
    per_code_ms = a.elapsed_time(b)
    same = all(torch.equal(bd.outputs[i].bits, per_code[i][0]) and
               torch.equal(bd.outputs[i].iterations, per_code[i][3]) for i in range(len(mx.codes)))
    # every replay of the demap graph wrote the eager LLRs
    demap_same = all(torch.equal(x, y) for x, y in zip(bd.inputs, eager))

    errors = 0
    failed = 0
    for g in mx.groups:
        bits = bd.outputs[g.code_index].bits[g.row0:g.row1]
        e = (bits != g.frames.info).sum(dim=1).cpu().numpy()
        errors += int(e.sum())
        failed += int((e > 0).sum())
    dec_bytes = 0.0
    for i, c in enumerate(mx.codes):
        dec_bytes += decoder_bytes(c, bd.outputs[i].iterations.cpu().numpy())
    demap_bytes = sum(g.frames.rx.numel() * (8 + 4 * int(np.log2(g.order))) for g in mx.groups)
    achieved = dec_bytes / (decode_ms * 1e-3) / 1e9
    value = args.steps * mx.info_bits / elapsed / 1e9
    return {
        "metric": "decoded info Gb/s at <=1e-5 post-FEC BER; p50/p99 frame latency",
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
        "data": "synthetic (seeded codes, uniform info bits, AWGN at per-frame Es/N0)",
        "config": {
            "workload": f"mixed: {frames} frames, codes {'/'.join(MIXED_CODES)} x Gray "
                        "4/16/64/256-QAM uniform, Es/N0 ~ U[wf-0.5, wf+1] dB per frame, "
                        "max-log LLRs, layered NMS 0.875, <=50 iterations, one batched launch",
            "frames": frames,
            "rows_per_code": mx.rows,
        },
        "frame_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                             "p99": round(float(np.percentile(step_ms, 99)), 3)},
        "post_fec_ber": errors / mx.info_bits,
        "fer": failed / frames,
        "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4),
                       "decode_per_code_launches": round(per_code_ms, 4)},
        "batch_equals_per_code": bool(same),
        "demap_graph_equals_eager": bool(demap_same),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "ldpc_batch (decode)",
            "bytes_per_launch": dec_bytes,
        },
        "demap_roofline": {
            "achieved": round(demap_bytes / (demap_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "bytes_per_launch": demap_bytes,
        },
    }


def run_shaped(args: argparse.Namespace) -> dict:
    """Shaped 256-QAM (Maxwell-Boltzmann nu = 0.05) by probabilistic
    amplitude shaping over the rate-5/6 code: 2 x 2^22 received symbols
    (3,106 whole codewords of 2,700 symbols and 2,408 filler symbols)
    demapped with exact LLRs and the prior, the codeword LLRs gathered out of
    the shaping bit layout, layered sum-product of at most 30 iterations.
    ``value`` is that tail's rate in the code's information bits (the C5
    metric); the CCDM dematcher that turns the decoded amplitude bits back
    into payload is timed after it (``pas``), with the payload rate of the
    whole chain."""
    import torch

    from commkit_amd import Constellation, SumProduct
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import llr_permuted
    from commkit_amd.operating import SHAPED_WATERFALL_DB
    from commkit_amd.workloads import make_pas_frames, seeded_code

    dev = torch.device("cuda", 0)
    name, nu = "r5/6-z360", 0.05
    code = seeded_code(name)
    c = Constellation.qam(256).shaped(nu=nu)
    esn0 = round(SHAPED_WATERFALL_DB[(name, 256, nu)] + 0.5, 3) if args.esn0 is None else args.esn0
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

    pas = fr.pas
    ccdm = pas.matcher

    perm32 = fr.perm.to(torch.int32).contiguous()

    def demap():
        # exact LLRs with the prior, written straight into the codewords' bit
        # layout (the filler symbols after the last codeword carry none)
        return llr_permuted(c, rx, nv, "exact", perm32, F)

    def dematch(fec_bits):
        payload, ok = ccdm._dematch_labels(fec_bits[:, : ccdm.n])
        return torch.cat([payload, fec_bits[:, ccdm.n:]], dim=1), ok

    for _ in range(args.warmup):
        out = _decode(code, demap(), method)
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    step_ms = []
    t0 = time.perf_counter()
    for s_ in range(args.steps):
        ts = time.perf_counter()
        ev[s_][0].record(stream)
        llr_c = demap()
        ev[s_][1].record(stream)
        out = _decode(code, llr_c, method)
        ev[s_][2].record(stream)
        ev[s_][2].synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    demap_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    decode_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    bits, _, valid, iters = out
    # The dematcher after the decoder, timed on its own.
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
    achieved = dec_bytes / (decode_ms * 1e-3) / 1e9
    served = shaped_served(args, code, c, fr, nv, F, method, dev)
    return {
        "metric": "decoded info Gb/s at <=1e-5 post-FEC BER; p50/p99 frame latency",
        "value": round(args.steps * F * code.k / elapsed / 1e9, 3),
        "unit": "Gb/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (uniform payload, CCDM amplitudes, code parity and payload on "
                "signs; AWGN)",
        "config": {
            "workload": f"shaped: 2 x 2^22 symbols of 256-QAM shaped(nu={nu}), {F} PAS "
                        f"codewords of {name} + {filler} filler symbols, exact LLRs with prior, "
                        "layered sum-product <=30 iterations, CCDM dematcher",
            "frames": F,

            "esn0_db": esn0,
            "code": name,
        },
        "frame_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                             "p99": round(float(np.percentile(step_ms, 99)), 3)},
        "post_fec_ber": fec_errors / (F * code.k),
        "fer": float((errors > 0).mean()),
        "mean_iterations": float(iters_h.mean()),
        "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4)},
        "pas": {
            "payload_bits_per_codeword": pas.k,
            "dematch_ms": round(dematch_ms, 4),
            "payload_ber": float(errors.sum() / (F * pas.k)),
            "frames_flagged": flagged,
            "payload_gbps_with_dematch": round(F * pas.k / (elapsed / args.steps
                                                            + dematch_ms * 1e-3) / 1e9, 3),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (PROFILED_TRAFFIC_SHAPED["bytes"] if F == PROFILED_TRAFFIC_SHAPED["frames"]
                        and args.esn0 is None else None),
            "kernel": "ldpc_qc_layered<sum-product> (decode)",
            "bytes_per_launch": dec_bytes,
        },
        "demap_roofline": {
            "achieved": round(demap_bytes / (demap_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "bytes_per_launch": demap_bytes,
            "note": "the demapper writes the LLRs straight into the codewords' bit layout "
                    "(ck_llr_separable_permuted)",
        },
        "served": served,
    }


def shaped_served(args, code, c, fr, nv: float, F: int, method, dev) -> dict:
    """The shaped frames at saturation inside the serving loop: requests of
    U{1..256} whole codewords, all queued at once; each request's symbols are
    demapped (exact, with the prior) straight into its ring slot, in codeword
    bit order, and decoded by continuous batching; decoded bits are copied
    out of the ring as requests complete, and one dematcher launch over all
    frames closes the run (its time is inside the measured wall clock: a
    dematcher launch costs about the same for one request as for all, one
    lane per codeword)."""
    import torch

    from commkit_amd.mapping.demap import llr_permuted
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

    def fill(slot, i):
        a, f = int(first[i]), int(sizes[i])
        llr_permuted(c, rx[a * S:(a + f) * S], nv, "exact", perm32, f, out=slot.llrs)

    # Results are copied out of the ring as requests complete (the slot is
    # reused after release), then scored once at the end.
    bits_out = torch.empty((F, code.k), dtype=torch.int8, device=dev)
    iters_out = torch.empty(F, dtype=torch.int32, device=dev)

    def on_done(slot, i):
        a, f = int(first[i]), int(sizes[i])
        bits_out[a:a + f].copy_(slot.bits)
        iters_out[a:a + f].copy_(slot.iterations)

    srv = Server([code], method, hbm_budget=16 << 30, max_inflight=args.max_inflight, device=dev)
    warm = Schedule(np.zeros(4), np.arange(4), np.zeros(4, np.int64), sizes[:4], np.zeros(4, np.int64))
    drive(srv, warm, fill)
    t0 = time.perf_counter()
    res = drive(srv, sched, fill, on_done=on_done)
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
    print(json.dumps({"progress": "shaped served", "gbps": round(res.gbps, 3)}), file=sys.stderr,
          flush=True)  # fmt: skip
    return {
        "gbps": round(res.gbps, 3),
        "payload_gbps_with_dematch": round(F * fr.pas.k / elapsed / 1e9, 3),
        "requests": int(n_req),
        "frames": int(F),
        "launches": res.launches,
        "max_inflight": args.max_inflight,
        "frame_latency_ms": {"p50": round(res.percentile(50), 3),
                             "p99": round(res.percentile(99), 3)},
        "payload_ber": errors / (F * fr.pas.k),
        "frames_flagged": int((~dm_ok).sum().item()),
        "mean_iterations": it_sum / F,
        "tail_hbm_frac": round(tail_bytes / res.elapsed_s / 1e9 / HBM_PEAK_GBS, 4),
        "note": "gbps: the code's information bits over the wall clock of the loop (host "
                "admission, per-request demap into the slot, batched decode); the payload "
                "rate adds the dematcher launch after the loop; algorithmic bytes of demap + "
                "decode over 8 TB/s",
    }


def run_serving(args: argparse.Namespace) -> dict:
    """32 Poisson request streams (stream s: code s mod 3, Gray QAM order
    4/16/64/256 by s // 3, request sizes uniform in 1..256 frames, Es/N0 a
    reflected random walk of 0.1 dB steps over [wf - 0.5, wf + 1]) served by
    the continuous-batching server under a 16 GiB HBM budget.  Requests carry
    channel LLRs, copied from pools of pre-generated frames per (code, QAM,
    Es/N0 step).  First the saturated throughput (every request queued at
    once), then each offered load as a fraction of it; per-frame latency is
    completion seen by the host minus arrival."""
    import torch

    from commkit_amd import Constellation, MinSum
    from commkit_amd.mapping.demap import llr_any
    from commkit_amd.operating import waterfall_db
    from commkit_amd.serve import Schedule, Server, burst_schedule, drive, poisson_schedule
    from commkit_amd.workloads import MIXED_CODES, MIXED_ORDERS, make_frames, seeded_code

    dev = torch.device("cuda", 0)
    streams, bins, max_frames = 32, 16, 256
    codes = [seeded_code(n) for n in MIXED_CODES]
    pairs = [(s % 3, MIXED_ORDERS[(s // 3) % 4]) for s in range(streams)]
    pools: dict = {}
    for ci, order in sorted(set(pairs)):
        c = Constellation.qam(order)
        wf = waterfall_db(MIXED_CODES[ci], order)
        per_bin = []
        for b in range(bins):
            fr = make_frames(codes[ci], c, wf - 0.5 + 0.1 * b, max_frames,
                             seed=100 * ci + order + 7 * b, device=dev)  # fmt: skip
            per_bin.append(llr_any(c, fr.rx, float(fr.noise_var[0]), "maxlog")
                           .reshape(max_frames, codes[ci].n))  # fmt: skip
        pools[(ci, order)] = per_bin
        print(json.dumps({"progress": "pool", "code": MIXED_CODES[ci], "order": order}),
              file=sys.stderr, flush=True)  # fmt: skip
    torch.cuda.synchronize()
    method = MinSum(max_iterations=50)
    srv = Server(codes, method, hbm_budget=16 << 30, max_inflight=args.max_inflight, device=dev)
    mean_bits = float(np.mean([codes[ci].k for ci, _ in pairs])) * (1 + max_frames) / 2

    def make_fill(sched: Schedule):
        def fill(slot, i):
            ci, order = pairs[int(sched.stream[i])]
            slot.llrs.copy_(pools[(ci, order)][int(sched.snr_bin[i])][: slot.frames])
        return fill

    def codes_of(sched: Schedule) -> Schedule:
        sched.code = np.asarray([pairs[int(s)][0] for s in sched.stream])
        return sched

    # Saturated: a burst of requests all present at the start.
    n_sat = 3000
    warm = codes_of(burst_schedule(64, streams=streams, num_codes=3, max_frames=max_frames,
                                   snr_bins=bins, seed=2))  # fmt: skip
    drive(srv, warm, make_fill(warm))
    burst = codes_of(burst_schedule(n_sat, streams=streams, num_codes=3, max_frames=max_frames,
                                    snr_bins=bins, seed=1))  # fmt: skip
    sat = drive(srv, burst, make_fill(burst))
    # Algorithmic bytes of the saturated run.  A frame's iteration count is
    # the same alone and in any batch, so the pools are decoded per call,
    # outside the timed region, for the counts.
    from commkit_amd.coding.ldpc import _decode

    iters_cum = {key: [np.cumsum(_decode(codes[key[0]], llr, method)[3].cpu().numpy())
                       for llr in per_bin] for key, per_bin in pools.items()}  # fmt: skip
    sat_bytes = 0.0
    for i in range(len(burst)):
        ci, order = pairs[int(burst.stream[i])]
        f = int(burst.frames[i])
        cd = codes[ci]
        its = int(iters_cum[(ci, order)][int(burst.snr_bin[i])][f - 1])
        sat_bytes += f * (8 * cd.n + cd.k + 5) + 8.0 * cd.checks.size * its
    print(json.dumps({"progress": "saturated", "gbps": round(sat.gbps, 3),
                      "launches": sat.launches}), file=sys.stderr, flush=True)  # fmt: skip
    loads = []
    fracs = [float(x) for x in args.loads.split(",") if x]
    if args.bound_load not in fracs:
        fracs = sorted(fracs + [args.bound_load])
    for frac in fracs:
        rate = frac * sat.gbps * 1e9 / mean_bits
        sched = codes_of(poisson_schedule(rate, args.serve_seconds, streams=streams, num_codes=3,
                                          max_frames=max_frames, seed=int(frac * 1000)))  # fmt: skip
        res = drive(srv, sched, make_fill(sched))
        loads.append({
            "offered_fraction": frac,
            "offered_gbps": round(frac * sat.gbps, 3),
            "requests": res.requests,
            "frames": res.frames,
            "carried_gbps": round(res.gbps, 3),
            "frame_latency_ms": {"p50": round(res.percentile(50), 3),
                                 "p99": round(res.percentile(99), 3)},
            "admissions_deferred": res.rejected,
            "launches": res.launches,
        })
        print(json.dumps({"progress": "load", **loads[-1]}), file=sys.stderr, flush=True)
    srv.close()
    top = loads[-1] if loads else None
    at = next(x for x in loads if x["offered_fraction"] == args.bound_load)
    bound = {"offered_fraction": args.bound_load, "offered_gbps": at["offered_gbps"],
             "p99_bound_ms": args.p99_bound_ms, "p99_ms": at["frame_latency_ms"]["p99"],
             "met": at["frame_latency_ms"]["p99"] <= args.p99_bound_ms}
    return {
        "metric": "decoded info Gb/s at <=1e-5 post-FEC BER; p50/p99 frame latency",
        "value": round(sat.gbps, 3),
        "unit": "Gb/s",
        "n_gpus": 1,
        "steps": sat.launches,
        "warmup": 1,
        "ms_per_step": round(sat.elapsed_s / max(sat.launches, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded codes, AWGN; request LLRs from pre-generated frame pools)",
        "config": {
            "workload": "serving: 32 Poisson streams, codes r5/6-z360/r2/3-z96/r1/2-z27, "
                        "Gray 4/16/64/256-QAM, requests U{1..256} frames, Es/N0 random walk "
                        "in [wf-0.5, wf+1] dB, 16 GiB HBM budget, layered NMS 0.875 <=50 "
                        "iterations, continuous batching",
            "saturation_requests": n_sat,
            "max_inflight": args.max_inflight,
            "serve_seconds": args.serve_seconds,
        },
        "latency_bound": bound,
        "frame_latency_ms": top["frame_latency_ms"] if top else None,
        "saturated": {"gbps": round(sat.gbps, 3), "requests": sat.requests,
                      "launches": sat.launches,
                      "decoder_hbm_frac": round(sat_bytes / sat.elapsed_s / 1e9 / HBM_PEAK_GBS, 4),
                      "frame_latency_ms": {"p50": round(sat.percentile(50), 3),
                                           "p99": round(sat.percentile(99), 3)}},
        "loads": loads,
    }


def main() -> int:
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(frames shard by whole codewords across GPUs; no multi-GPU run here).",
              file=sys.stderr)  # fmt: skip
        return 2
    import torch

    from commkit_amd import Constellation, MinSum, metrics
    from commkit_amd.coding.ldpc import _decode
    from commkit_amd.mapping.demap import llr_any
    from commkit_amd.operating import waterfall_db
    from commkit_amd.workloads import make_frames, seeded_code

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        copy_calibration(dev)
        return 0
    if args.workload == "mixed":
        print(json.dumps(run_mixed(args)), flush=True)
        return 0
    if args.workload == "shaped":
        print(json.dumps(run_shaped(args)), flush=True)
        return 0
    if args.workload == "serving":
        print(json.dumps(run_serving(args)), flush=True)
        return 0
    code = seeded_code(CODE)
    c = Constellation.qam(ORDER)
    esn0 = args.esn0 if args.esn0 is not None else round(waterfall_db(CODE, ORDER) + MARGIN_DB, 3)
    F = args.frames
    fr = make_frames(code, c, esn0, F, seed=0, device=dev)
    nv = float(fr.noise_var[0])
    rx = fr.rx.contiguous()
    method = MinSum(max_iterations=50)
    torch.cuda.synchronize()

    def step():
        llr = llr_any(c, rx, nv, "maxlog")
        return _decode(code, llr.reshape(F, code.n), method)

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    step_ms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        ev[s][0].record(stream)
        llr = llr_any(c, rx, nv, "maxlog")
        ev[s][1].record(stream)
        ev[s][2].record(stream)
        out = _decode(code, llr.reshape(F, code.n), method)
        ev[s][3].record(stream)
        ev[s][3].synchronize()
        step_ms.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    demap_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    decode_ms = float(np.mean([e[2].elapsed_time(e[3]) for e in ev]))

    bits, app, valid, iters = out
    iters_h = iters.cpu().numpy()
    # Scored by the library's own kernels, after the clock stopped.
    ber = metrics.ber(bits.reshape(-1), fr.info.reshape(-1))
    fer = metrics.fer(bits.reshape(-1), fr.info.reshape(-1), info_bits=code.k)
    info_bits = F * code.k
    value = args.steps * info_bits / elapsed / 1e9
    dec_bytes = decoder_bytes(code, iters_h)
    achieved = dec_bytes / (decode_ms * 1e-3) / 1e9
    demap_bytes = rx.numel() * (8 + 4 * c.bits_per_symbol)
    line = {
        "metric": "decoded info Gb/s at <=1e-5 post-FEC BER; p50/p99 frame latency",
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
        },
        "frame_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                             "p99": round(float(np.percentile(step_ms, 99)), 3)},
        "post_fec_ber": ber,
        "fer": fer,
        "mean_iterations": float(iters_h.mean()),
        "kernels_ms": {"demap": round(demap_ms, 4), "decode": round(decode_ms, 4)},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (PROFILED_TRAFFIC["bytes"] if F == PROFILED_TRAFFIC["frames"]
                        and args.esn0 is None else None),
            "kernel": "ldpc_qc_layered (decode)",
            "bytes_per_launch": dec_bytes,
        },
        "demap_roofline": {
            "achieved": round(demap_bytes / (demap_ms * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "bytes_per_launch": demap_bytes,
        },
        "tail_roofline": {
            "achieved": round((demap_bytes + dec_bytes) / ((demap_ms + decode_ms) * 1e-3) / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
        },
        "hbm_copy_GBs": round(copy_bandwidth(dev), 1),
    }
    line["reference_benchmark"] = reference_benchmark(dev)
    line["reference_llr_benchmarks"] = reference_llr_benchmarks(dev)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(code, c, fr.rx.cpu().numpy(), fr.info.cpu().numpy(),
                                            nv, args.cpu_seconds)  # fmt: skip
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
