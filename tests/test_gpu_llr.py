"""HIP demapper against the FP64 definition and the float32 CPU oracle (GPU).

Tolerance (the LLR contract, stated here): with L64 the float64 LLR of the
definition (float64 constellation, float64 arithmetic, the same complex64
symbols), a GPU LLR L32 meets

- max-log: ``|L32 - L64| <= 1e-5 * |L64|`` on every element, with no floor
  (the kernels form metric differences in float64 from float64 tables and
  a float64 ``1/sigma^2``, so only the float32 output rounds);
- exact (log-sum-exp): ``|L32 - L64| <= 1e-5 * max(|L64|, 2^-5)``: relative
  1e-5 except within 2^-5 of zero, where the float32 correction sums
  ``log(1 + T)`` allow an absolute 3e-7.

The sign (hard decision) must agree wherever |L64| exceeds the floor.  Each
case prints its largest relative error and how many elements fell under
the floor.
"""

import numpy as np
import pytest

from commkit_amd import Constellation, Signal, compute_llr
from commkit_amd.mapping.demap import log_prior, opposite_levels
from commkit_amd.mapping.constellation import pack_bits
from oracle import cpu
from oracle.llr64 import llr_fp64, llr_tolerance_ok

FLOOR = {"maxlog": 0.0, "exact": 2.0**-5}

pytestmark = pytest.mark.gpu

CASES = {
    "qpsk": (lambda: Constellation.qam(4), 0.3),
    "16qam": (lambda: Constellation.qam(16), 10 ** (-0.8)),
    "64qam": (lambda: Constellation.qam(64), 10 ** (-1.4)),
    "256qam": (lambda: Constellation.qam(256), 10 ** (-2.0)),
    "64qam-shaped": (lambda: Constellation.qam(64).shaped(nu=0.05), 10 ** (-1.3)),
    "256qam-shaped": (lambda: Constellation.qam(256).shaped(nu=0.05), 10 ** (-1.8)),
    "8psk": (lambda: Constellation.psk(8), 0.05),
    "8qam": (lambda: Constellation.qam(8), 0.05),
    "32qam-cross": (lambda: Constellation.qam(32), 10 ** (-1.2)),
    "128qam-cross": (lambda: Constellation.qam(128), 10 ** (-1.7)),
    "1024qam": (lambda: Constellation.qam(1024), 10 ** (-2.8)),
}


def _noisy(c, n, nv, seed):
    r = np.random.default_rng(seed)
    if c.pmf is not None:
        idx = r.choice(c.order, n, p=c.pmf)
    else:
        idx = r.integers(0, c.order, n)
    noise = np.sqrt(nv / 2) * (r.standard_normal(n) + 1j * r.standard_normal(n))
    return (c.points[idx] + noise).astype(np.complex64)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("method", ["maxlog", "exact"])
def test_llr_against_fp64(gpu, name, method):
    import torch

    make, nv = CASES[name]
    c = make()
    x = _noisy(c, 20000, nv, seed=11)
    got = compute_llr(torch.from_numpy(x).cuda(), noise_var=nv, constellation=c, method=method)
    assert got.dtype == torch.float32 and got.shape == (x.size * c.bits_per_symbol,)
    got = got.cpu().numpy()
    ref = llr_fp64(x, c.points, c.bit_labels, c.pmf, nv, method)
    ok, under = llr_tolerance_ok(got, ref, floor=FLOOR[method])
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
    err_abs = np.abs(got - ref)[under].max() if under.any() else 0.0
    print(f"{name} {method}: max rel err above floor {rel[~under].max():.2e}, "
          f"{int(under.sum())} of {ref.size} under the floor (max abs err {err_abs:.2e})")
    assert ok, f"max rel err {rel[~under].max():.3e}, abs under floor {err_abs:.3e}"
    assert np.array_equal(np.sign(got[~under]), np.sign(ref[~under]))


@pytest.mark.parametrize("order", [16, 64, 256])
def test_llr_against_c_oracle(gpu, order):
    import torch

    c = Constellation.qam(order)
    nv = 0.02
    x = _noisy(c, 5000, nv, seed=12)
    re, im = c._axes
    inv = float(np.float32(1 / nv))
    closed = []
    for a in (re, im):
        lo, hi = opposite_levels(a.labels)
        closed.append((a.levels, pack_bits(a.labels), lo, hi, a.bits))
    want = cpu.llr_separable_maxlog(x, closed, inv)
    got = compute_llr(torch.from_numpy(x).cuda(), noise_var=nv, constellation=c).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-3)
    sep = [(a.levels, a.labels, log_prior(a.pmf, a.levels.size), a.bits) for a in (re, im)]
    want = cpu.llr_separable(x, sep, inv, True)
    got = compute_llr(torch.from_numpy(x).cuda(), noise_var=nv, constellation=c,
                      method="exact").cpu().numpy()  # fmt: skip
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-3)


def test_noiseless_signs_at_scale(gpu):
    import torch

    c = Constellation.qam(64)
    r = np.random.default_rng(13)
    n = 1 << 22
    idx = torch.from_numpy(r.integers(0, 64, n)).cuda()
    pts = torch.from_numpy(c.points.astype(np.complex64)).cuda()
    x = pts[idx]
    for method in ("maxlog", "exact"):
        llr = compute_llr(x, noise_var=0.05, constellation=c, method=method)
        labels = torch.from_numpy(c.bit_labels).cuda()[idx].reshape(-1)
        assert torch.equal((llr < 0).to(torch.int8), labels)


def test_contract(gpu):
    import torch

    c = Constellation.qam(16)
    x = _noisy(c, 15, 0.1, seed=14).reshape(3, 5)
    host = compute_llr(x, noise_var=0.1, constellation=c)
    assert isinstance(host, np.ndarray) and host.shape == (3, 20) and host.dtype == np.float32
    dev = compute_llr(torch.from_numpy(x).cuda(), noise_var=0.1, constellation=c)
    assert dev.device.type == "cuda" and dev.shape == (3, 20)
    np.testing.assert_array_equal(dev.cpu().numpy(), host)
    sig = compute_llr(Signal(x, constellation=c), esn0_db=10.0)
    np.testing.assert_array_equal(sig, host)
    via = c.llr(x, noise_var=0.1)
    np.testing.assert_array_equal(via, host)
    empty = compute_llr(np.zeros(0, np.complex64), noise_var=0.1, constellation=c)
    assert empty.shape == (0,)
    with pytest.raises(ValueError, match="sps"):
        compute_llr(Signal(x, constellation=c, sps=2), noise_var=0.1)
    from commkit_amd import PrecisionWarning

    with pytest.warns(PrecisionWarning):
        compute_llr(x.astype(np.complex128), noise_var=0.1, constellation=c)


def test_real_pam(gpu):
    c = Constellation.pam(4)
    r = np.random.default_rng(15)
    x = (c.points[r.integers(0, 4, 3000)] + 0.2 * r.standard_normal(3000)).astype(np.float32)
    nv = 0.04  # real noise variance
    got = compute_llr(x, noise_var=nv, constellation=c)
    ref = llr_fp64(x, c.points, c.bit_labels, None, nv, "maxlog")
    ok, _ = llr_tolerance_ok(got, ref)
    assert ok


@pytest.mark.parametrize("make,method", [
    (lambda: Constellation.qam(4), "maxlog"),
    (lambda: Constellation.qam(64), "maxlog"),
    (lambda: Constellation.qam(16), "exact"),
    (lambda: Constellation.qam(256).shaped(nu=0.05), "exact"),
    (lambda: Constellation.qam(64).shaped(nu=0.05), "maxlog"),
], ids=["qpsk-maxlog", "64qam-maxlog", "16qam-exact", "256qam-shaped-exact", "64qam-shaped-maxlog"])
def test_llr_rows_equal_per_row_calls(gpu, make, method):
    """Per-row noise (one launch) gives each row the LLRs of a scalar call, bit for bit."""
    import torch

    from commkit_amd.mapping.demap import llr_any, llr_rows

    c = make()
    rng = np.random.default_rng(11)
    R, S = 9, 301
    pts = c.points[rng.integers(0, c.order, size=(R, S))]
    nv = rng.uniform(0.005, 0.4, size=R)
    x = pts + np.sqrt(nv[:, None] / 2) * (rng.standard_normal((R, S)) + 1j * rng.standard_normal((R, S)))
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    got = llr_rows(c, xt, nv, method)
    out = torch.full((R, S * c.bits_per_symbol), float("nan"), device="cuda")
    assert llr_rows(c, xt, nv, method, out=out) is out
    for r in range(R):
        want = llr_any(c, xt[r], float(nv[r]), method)
        assert torch.equal(got[r], want), f"row {r}"
        assert torch.equal(out[r], want), f"row {r} (out=)"
    with pytest.raises(ValueError):
        llr_rows(c, xt, nv[:-1], method)
    with pytest.raises(ValueError):
        llr_rows(Constellation.psk(8), xt, nv, method)


def test_llr_rows_inverse_variances_in_a_graph(gpu):
    """inv_s2 on the device gives the LLRs of noise_var; a captured call
    replays to the eager result every time, with eager work between the
    replays."""
    import torch

    from commkit_amd.mapping.demap import llr_rows

    c = Constellation.qam(64)
    rng = np.random.default_rng(12)
    R, S = 7, 513
    nv = rng.uniform(0.01, 0.2, size=R)
    x = c.points[rng.integers(0, 64, size=(R, S))] + 0.1 * rng.standard_normal((R, S))
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    want = llr_rows(c, xt, nv, "maxlog")
    inv = torch.from_numpy(1.0 / np.maximum(nv, 1e-20)).cuda()
    out = torch.zeros_like(want)
    assert torch.equal(llr_rows(c, xt, None, "maxlog", out=out, inv_s2=inv), want)
    graph = torch.cuda.CUDAGraph()
    out.zero_()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        llr_rows(c, xt, None, "maxlog", out=out, inv_s2=inv)
    for _ in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, want)
        llr_rows(c, xt, nv, "maxlog")  # eager work between replays
    with pytest.raises(ValueError, match="inv_s2"):
        llr_rows(c, xt, None, "maxlog", inv_s2=inv.float())


@pytest.mark.parametrize(("make", "method", "period", "blocks"), [
    (lambda: Constellation.qam(256).shaped(nu=0.05), "exact", 21600, 3),
    (lambda: Constellation.qam(16), "maxlog", 1000, 7),
    (lambda: Constellation.qam(64), "maxlog", 66, 40),
], ids=["256qam-shaped-exact-pas", "16qam-closed-ragged", "64qam-small-period"])
def test_llr_permuted_equals_demap_then_gather(gpu, make, method, period, blocks):
    """The permuted store gives the gathered LLRs bit for bit, for blocks that
    end inside a symbol and blocks shorter than a tile; symbols past the
    blocks are left alone."""
    import torch

    from commkit_amd.mapping.demap import llr_any, llr_permuted

    c = make()
    k = c.bits_per_symbol
    rng = np.random.default_rng(period)
    nsym = -(-blocks * period // k) + 5
    x = _noisy(c, nsym, 0.02, seed=period)
    xt = torch.from_numpy(x).cuda()
    if period == 21600:
        from commkit_amd.coding import PAS
        from commkit_amd.workloads import seeded_code

        perm = PAS(seeded_code("r5/6-z360"), c)._tables(xt.device)[0]
    else:
        perm = torch.from_numpy(rng.permutation(period)).cuda()
    iperm = torch.argsort(perm)
    want = llr_any(c, xt, 0.02, method)[: blocks * period].view(blocks, period)[:, iperm]
    got = llr_permuted(c, xt, 0.02, method, perm.to(torch.int32).contiguous(), blocks)
    assert torch.equal(got, want)
    with pytest.raises(ValueError, match="int32"):
        llr_permuted(c, xt, 0.02, method, perm, blocks)
    with pytest.raises(ValueError, match="symbols"):
        llr_permuted(c, xt, 0.02, method, perm.to(torch.int32), blocks + 100)


def _natural_and_bimodal():
    """16-QAM variants the unimodal Gray form must not take: natural-binary
    axis labels (not Gray), and a separable prior whose per-axis log is
    convex (0.4, 0.1, 0.1, 0.4 over the levels), so the metric along an axis
    can have two peaks."""
    from dataclasses import replace

    c = Constellation.qam(16)
    lev = np.unique(np.round(c.points.real, 12))
    ii = np.searchsorted(lev, np.round(c.points.real, 12))
    qq = np.searchsorted(lev, np.round(c.points.imag, 12))
    lab = np.stack([(ii >> 1) & 1, ii & 1, (qq >> 1) & 1, qq & 1], 1).astype(np.int8)
    p1 = np.array([0.4, 0.1, 0.1, 0.4])
    pmf = p1[ii] * p1[qq]
    return replace(c, bit_labels=lab), replace(c, pmf=pmf / pmf.sum())


GRAY_SWEEP = {
    "16qam": (lambda: Constellation.qam(16), [-3.0, 5.0, 10.0, 20.0, 30.0, 40.0]),
    "64qam": (lambda: Constellation.qam(64), [-3.0, 3.0, 10.0, 15.0, 20.0, 30.0, 40.0]),
    "256qam": (lambda: Constellation.qam(256), [-3.0, 3.0, 8.0, 15.0, 20.0, 30.0, 40.0]),
    "256qam-shaped": (lambda: Constellation.qam(256).shaped(nu=0.05),
                      [-3.0, 0.0, 3.0, 8.0, 13.7, 20.0, 30.0]),
    "1024qam": (lambda: Constellation.qam(1024), [-3.0, 3.0, 8.0, 13.7, 25.0, 30.0, 40.0]),
    "16qam-natural": (lambda: _natural_and_bimodal()[0], [-3.0, 5.0, 15.0, 30.0]),
    "16qam-bimodal-prior": (lambda: _natural_and_bimodal()[1], [-3.0, 5.0, 15.0, 30.0]),
    "16psk": (lambda: Constellation.psk(16), [-3.0, 3.0, 10.0, 20.0, 30.0]),
    "32qam-cross": (lambda: Constellation.qam(32), [-3.0, 3.0, 10.0, 20.0, 30.0]),
    "128qam-cross": (lambda: Constellation.qam(128), [-3.0, 3.0, 10.0, 20.0, 30.0]),
}


@pytest.mark.parametrize("name", list(GRAY_SWEEP))
@pytest.mark.parametrize("method", ["maxlog", "exact"])
def test_llr_sweep_against_fp64(gpu, name, method):
    """Level and point metrics over an Es/N0 sweep: Gray axes with a concave
    (uniform or Maxwell-Boltzmann) log prior take the unimodal Gray forms,
    whose far-under halves (high Es/N0) are summed relative to their own best
    level; the natural-label and two-peak-prior cases take the general form;
    PSK and cross QAM the point form.  Same tolerance as above on every
    element."""
    import torch

    make, dbs = GRAY_SWEEP[name]
    c = make()
    if method == "maxlog" and c.pmf is None and c._axes is not None and name != "16qam-natural":
        pytest.skip("uniform Gray max-log takes the closed form, which "
                    "test_llr_against_fp64 holds to FP64 for QPSK to 1024-QAM")  # fmt: skip
    for db in dbs:
        nv = 10 ** (-db / 10)
        x = _noisy(c, 4096 if c.order >= 1024 else 8192, nv, seed=int(db * 10) + 107)
        got = compute_llr(torch.from_numpy(x).cuda(), noise_var=nv, constellation=c,
                          method=method).cpu().numpy()  # fmt: skip
        ref = llr_fp64(x, c.points, c.bit_labels, c.pmf, nv, method)
        ok, under = llr_tolerance_ok(got, ref, floor=FLOOR[method])
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), max(FLOOR[method], 1e-30))
        print(f"{name} {method} {db} dB: worst error / bar {rel.max() / 1e-5:.3f}")
        assert ok, f"{db} dB: worst error {rel.max() / 1e-5:.3f} x the bar"
        assert np.array_equal(np.sign(got[~under]), np.sign(ref[~under]))


def test_prior_steeper_than_the_likelihood(gpu):
    """A log prior far steeper than the likelihood (-1000 (i - 3/2)^2 nats
    over 16-QAM's axis levels: a symbol snapped to an edge level has a level
    about 2000 nats above it, beyond the float64 range of the half sums)
    takes the second pass with the peak as reference.  Such a prior lies
    below the Python API's 1e-40 clip, so the tables go through the C-ABI;
    the expected LLRs are the definition in float64 over the 16 points."""
    import ctypes

    import torch

    from commkit_amd import _native
    from commkit_amd.mapping.demap import CkAxis, _axis_tables
    from commkit_amd.mapping.lattice import lattice_coordinates

    c = Constellation.qam(16)
    lp = -1000.0 * (np.arange(4) - 1.5) ** 2
    keep, axes = [], []
    for a in c._axes:
        t = list(_axis_tables(a))
        t[1] = lp.copy()
        t = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in t]
        keep.append(t)
        axes.append(CkAxis(*(v.data_ptr() for v in t), int(t[0].numel()), int(t[5].numel())))
    nv = 1.0
    x = _noisy(c, 4096, nv, seed=11)
    x[:64] = 1.2 * np.exp(2j * np.pi * np.arange(64) / 64)  # snapped to edge levels
    xt = torch.from_numpy(x.astype(np.complex64)).cuda()
    out = torch.empty((x.size, c.bits_per_symbol), dtype=torch.float32, device="cuda")
    lib = _native.lib()
    _native.check(lib.ck_llr_separable(
        xt.data_ptr(), x.size, ctypes.byref(axes[0]), ctypes.byref(axes[1]),
        ctypes.c_double(1.0 / nv), 1, out.data_ptr(), _native.stream_of(xt.device)))  # fmt: skip
    got = out.cpu().numpy().reshape(-1)
    side, where = lattice_coordinates(c.points)
    lp_pt = np.empty(c.order)
    for i in range(side):
        for q in range(side):
            lp_pt[where[i, q]] = lp[i] + lp[q]
    xs = x.astype(np.complex64).astype(np.complex128)
    metric = -np.abs(xs[:, None] - c.points[None, :]) ** 2 / nv + lp_pt[None, :]
    ref = np.stack([np.logaddexp.reduce(metric[:, c.bit_labels[:, b] == 0], axis=1) -
                    np.logaddexp.reduce(metric[:, c.bit_labels[:, b] == 1], axis=1)
                    for b in range(c.bits_per_symbol)], axis=1).reshape(-1)  # fmt: skip
    assert np.isfinite(got).all()
    ok, _ = llr_tolerance_ok(got, ref, floor=FLOOR["exact"])
    assert ok, f"worst error {np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 2**-5)):.3e}"


@pytest.mark.parametrize("make,method", [
    (lambda: Constellation.qam(64), "maxlog"),
    (lambda: Constellation.qam(16), "exact"),
    (lambda: Constellation.qam(256).shaped(nu=0.05), "exact"),
], ids=["64qam-maxlog", "16qam-exact", "256qam-shaped-exact"])  # fmt: skip
def test_segments_equal_per_request_calls(gpu, make, method):
    """Twenty requests through a SegmentsDemapper (the 16th launches the
    first 16, flush the rest): scalar noise, per-row noise and a code's bit
    layout, sizes from
    nothing to several tiles; each request's LLRs equal its own call bit for
    bit, and nothing outside the requests' outputs is written."""
    import torch

    from commkit_amd.mapping.demap import RowsDemapper, SegmentsDemapper, llr_any, llr_permuted

    c = make()
    k = c.bits_per_symbol
    rng = np.random.default_rng(21)
    dm = SegmentsDemapper(c, method, "cuda")
    rows_dm = RowsDemapper(c, method, "cuda")
    period = 36 * k // 2  # blocks end inside a symbol when k is odd per block
    perm = torch.from_numpy(rng.permutation(period).astype(np.int32)).cuda()
    iperm = torch.argsort(perm.long())
    checks = []
    for j in range(20):
        kind = j % 3
        if kind == 0:
            n = [0, 1, 255, 256, 257, 1000, 3001][j % 7]
            x = torch.from_numpy(_noisy(c, n, 0.03, seed=j)).cuda()
            out = torch.full((n, k), float("nan"), device="cuda")
            dm.add(x, out, noise_var=0.03 + 0.001 * j)
            checks.append((out, llr_any(c, x, 0.03 + 0.001 * j, method).reshape(n, k)))
        elif kind == 1:
            R, S = 1 + j % 4, 97 + 13 * j
            x = torch.from_numpy(_noisy(c, R * S, 0.05, seed=j).reshape(R, S)).cuda()
            inv = torch.from_numpy(1.0 / rng.uniform(0.01, 0.2, size=R)).cuda()
            out = torch.full((R, S * k), float("nan"), device="cuda")
            dm.add(x, out, inv_s2=inv)
            checks.append((out, rows_dm(x, inv, torch.empty_like(out))))
        else:
            blocks = 1 + j % 5
            x = torch.from_numpy(_noisy(c, -(-blocks * period // k) + 3, 0.02, seed=j)).cuda()
            out = torch.full((blocks, period), float("nan"), device="cuda")
            dm.add(x, out, noise_var=0.02, perm=perm)
            want = llr_permuted(c, x, 0.02, method, perm, blocks)
            checks.append((out, want))
            flat = llr_any(c, x, 0.02, method)[: blocks * period].view(blocks, period)[:, iperm]
            assert torch.equal(want, flat)
    assert len(dm) == 4  # the 16th request launched the first 16
    dm.flush()
    assert len(dm) == 0
    for j, (got, want) in enumerate(checks):
        assert torch.equal(got, want), f"request {j}"
    with pytest.raises(ValueError, match="exactly one"):
        dm.add(x, out, perm=perm)


def _maxlog_factored(x: np.ndarray, points: np.ndarray, bit_labels: np.ndarray,
                     noise_var: float) -> np.ndarray:
    """Max-log LLRs of the definition with each half's best point located by
    float64 metrics and the difference taken in factored form,
    ``((pa - pc) (2x - pa - pc)).real-and-imag / sigma^2``: free of the
    cancellation of a difference of two large float64 maxima when x sits on
    a decision boundary (uniform prior)."""
    x = np.asarray(x).ravel().astype(np.complex128)
    pts = np.asarray(points, dtype=np.complex128)
    k = bit_labels.shape[1]
    halves = [(np.where(bit_labels[:, b] == 0)[0], np.where(bit_labels[:, b] == 1)[0])
              for b in range(k)]  # fmt: skip
    out = np.empty((x.size, k))
    for i0 in range(0, x.size, 256):
        xs = x[i0 : i0 + 256]
        d2 = (xs.real[:, None] - pts.real[None, :]) ** 2 + (xs.imag[:, None] - pts.imag[None, :]) ** 2
        for b, (zero, one) in enumerate(halves):
            pa = pts[zero[np.argmin(d2[:, zero], axis=1)]]
            pc = pts[one[np.argmin(d2[:, one], axis=1)]]
            num = ((pa.real - pc.real) * ((xs.real - pa.real) + (xs.real - pc.real)) +
                   (pa.imag - pc.imag) * ((xs.imag - pa.imag) + (xs.imag - pc.imag)))  # fmt: skip
            out[i0 : i0 + 256, b] = num / noise_var
    return out.reshape(-1)


@pytest.mark.parametrize("order", [16, 64, 256, 1024, 4096, 16384])
def test_closed_form_at_cell_boundaries(gpu, order):
    """The max-log closed form on symbols at every half-grid point of the
    axes (each level and each midpoint of neighbouring levels, rounded to
    float32) and at their float32 neighbours, where one float64 comparison
    decides the cell, plus noisy symbols; every element within 1e-5
    relative of the definition.  Up to 4096-QAM (6 bits per axis) the
    kernels use per-cell tables; 16384-QAM takes the snap-and-compare
    form."""
    import torch

    c = Constellation.qam(order)
    lev = c._axes[0].levels.astype(np.float64)
    half = np.sort(np.concatenate([lev, 0.5 * (lev[1:] + lev[:-1])])).astype(np.float32)
    grid = np.unique(np.concatenate([half, np.nextafter(half, np.float32(np.inf)),
                                     np.nextafter(half, np.float32(-np.inf))]))  # fmt: skip
    r = np.random.default_rng(order)
    other = r.uniform(lev[0] * 1.1, lev[-1] * 1.1, grid.size).astype(np.float32)
    nv = float((lev[1] - lev[0]) ** 2)  # one level step per standard deviation
    x = np.concatenate([grid + 1j * other, other + 1j * grid,
                        _noisy(c, 1024, nv, seed=order)]).astype(np.complex64)  # fmt: skip
    got = compute_llr(torch.from_numpy(x).cuda(), noise_var=nv, constellation=c).cpu().numpy()
    ref = _maxlog_factored(x, c.points, c.bit_labels, nv)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-37)
    rel[np.abs(ref) < 1e-37] = 0.0  # float32 denormal results
    print(f"{order}-QAM: {x.size} symbols, worst error / bar {rel.max() / 1e-5:.4f}")
    assert rel.max() <= 1e-5
    assert np.array_equal(np.sign(got[np.abs(ref) >= 1e-37]), np.sign(ref[np.abs(ref) >= 1e-37]))


def test_points_with_repeated_labels(gpu):
    """ck_llr_points with labels that are not a permutation (two points share
    a label, one label is absent): the kernel stages the tables as given,
    not in label order, and its LLRs still follow the definition with those
    labels (every element within 1e-5 relative, max-log and exact)."""
    import ctypes

    import torch

    from commkit_amd import _native

    c = Constellation.psk(8)
    labels = np.array([0, 1, 3, 2, 6, 7, 5, 0], dtype=np.int32)  # 4 absent, 0 twice
    bits = ((labels[:, None] >> np.arange(2, -1, -1)[None, :]) & 1).astype(np.int8)
    nv = 0.05
    x = _noisy(c, 8192, nv, seed=21)
    dev = torch.device("cuda", 0)
    xs = torch.from_numpy(x).to(dev)
    pts = torch.from_numpy(c.points.astype(np.complex128)).to(dev)
    lp = torch.zeros(8, dtype=torch.float64, device=dev)
    lab = torch.from_numpy(labels).to(dev)
    lib = _native.lib()
    for method in ("maxlog", "exact"):
        out = torch.empty((x.size, 3), dtype=torch.float32, device=dev)
        _native.check(lib.ck_llr_points(xs.data_ptr(), x.size, pts.data_ptr(), lp.data_ptr(),
                                        lab.data_ptr(), 8, 3, ctypes.c_double(1.0 / nv),
                                        int(method == "exact"), out.data_ptr(),
                                        _native.stream_of(dev)))  # fmt: skip
        got = out.cpu().numpy().reshape(-1)
        ref = llr_fp64(x, c.points, bits, None, nv, method)
        ok, under = llr_tolerance_ok(got, ref, floor=FLOOR[method])
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), max(FLOOR[method], 1e-30))
        print(f"repeated labels {method}: worst error / bar {rel.max() / 1e-5:.3f}")
        assert ok
