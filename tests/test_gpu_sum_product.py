"""Sum-product decoding against the CPU oracle on the shaped workload (GPU).

The workload is the shaped 256-QAM tail: probabilistic-amplitude-shaping
codewords of the rate-5/6 code (Z = 360, n = 21600), Maxwell-Boltzmann
nu = 0.05, exact LLRs with the prior written in codeword order, layered
SumProduct(max_iterations=30) with early stop, at the operating point
(shaped waterfall + 0.5 dB = 13.7 dB) and at the waterfall itself (13.2 dB,
where frames run about twice as long).

The GPU evaluates phi(x) = -log tanh(x/2) from the hardware exp2 /
reciprocal / log2; the oracle with glibc's logf / tanhf, as the reference's
Numba loop does (commkit/coding/_numba_loops.py:31-40).  The two agree to
rounding, not bit for bit, and what the bar can be follows from the float32
algorithm itself: where one edge dominates a row's phi total, total -
phi(|Q_e|) cancels to a few ulps of the total, and each ulp moves that
edge's message by about ln 2.  The oracle run with phi rounded differently
and nothing else (``cpu.ldpc_layered_probe``: phi in float64, rounded once)
moves its APP LLRs by up to about 0.9 at this workload, so no float32
restatement can meet the reference test's rtol/atol 1e-3
(tests/coding/test_ldpc_kernel.py:37-45, at 5 iterations of small LLRs:
``test_gpu_ldpc.test_sum_product_matches`` holds the GPU to it there).
The bar here:

- hard decisions of all n bits, iteration counts and valid flags identical
  on every frame, except frames that are listed and traced: decoded again
  on both sides with max_iterations = 1, 2, ... until the hard decisions
  first differ, where every flipped bit must be a near tie,
  |APP_gpu - APP_oracle| <= NEAR_TIE (the two have opposite signs, so both
  lie within NEAR_TIE of zero).  The GPU runs whole iterations, so the trace
  names the iteration and the bits, not the row inside it;
- on the other frames, every bit whose oracle APP is at most SMALL_APP in
  magnitude (no saturated message behind it) within rtol/atol 1e-3;
- and every quantile of |APP_gpu - APP_oracle| (median to maximum) within
  CONDITIONING times the same quantile for the probe, plus 1e-4: the GPU
  moves no further from the oracle than a change of phi's rounding does.

Every GPU kernel and tile must give the same outputs bit for bit (the lanes
of a row add its phi values in ascending edge order whatever their number).
"""

from dataclasses import replace

import numpy as np
import pytest

from commkit_amd import Constellation, SumProduct
from commkit_amd.workloads import seeded_code

from oracle import cpu

from .test_gpu_ldpc import gpu_decode, oracle_decode

pytestmark = pytest.mark.gpu

RTOL = ATOL = 1e-3
NEAR_TIE = 1e-2
SMALL_APP = 2.0
CONDITIONING = 4.0
QUANTILES = (0.5, 0.9, 0.99, 0.999, 1.0)
FRAMES = 64
CODE = "r5/6-z360"
NU = 0.05
# (kernel, frames per group, lanes per row, state in LDS): the QC kernel with
# one and two lanes per row (the tuned tile), two frames per group, and the
# generic CSR kernel.
SP_TILES = [(2, 1, 1, 0), (2, 1, 2, 0), (2, 2, 1, 0), (1, 0, 0, -1)]


@pytest.fixture(scope="module")
def shaped(gpu):
    """Exact codeword-order LLRs of FRAMES PAS codewords at each Es/N0."""
    import torch

    from commkit_amd.mapping.demap import llr_permuted
    from commkit_amd.operating import SHAPED_WATERFALL_DB
    from commkit_amd.workloads import make_pas_frames

    code = seeded_code(CODE)
    c = Constellation.qam(256).shaped(nu=NU)
    wf = SHAPED_WATERFALL_DB[(CODE, 256, NU)]
    out = {}
    for esn0 in (round(wf + 0.5, 3), wf):
        fr = make_pas_frames(code, c, esn0, FRAMES, seed=21, device=gpu)
        perm = fr.perm.to(torch.int32).contiguous()
        llr = llr_permuted(c, fr.rx.reshape(-1), float(fr.noise_var[0]), "exact", perm, FRAMES)
        torch.cuda.synchronize()
        out[esn0] = llr.cpu().numpy()
    return code, out


def probe_decode(code, llrs, method):
    """The oracle with phi rounded once from float64 (conditioning probe)."""
    full = np.zeros((llrs.shape[0], code.num_variables), np.float32)
    full[:, : code.n] = llrs
    row_ptr, cols, _ = code._schedule
    app, iters, valid = cpu.ldpc_layered_probe(full, row_ptr, cols,
                                               max_iterations=method.max_iterations,
                                               early_stop=method.early_stop)  # fmt: skip
    return app[:, : code.n], iters, valid


def trace(code, llr, method, tiles):
    """First iteration at which the hard decisions of one frame differ
    between the GPU (``tiles``) and the oracle: (iteration, flipped bits,
    GPU APP, oracle APP), or None."""
    for t in range(1, method.max_iterations + 1):
        m = replace(method, max_iterations=t)
        g = gpu_decode(code, llr[None], m, tiles)
        o = oracle_decode(code, llr[None], m)
        flip = np.flatnonzero((g[1][0] < 0) != (o[1][0] < 0))
        if flip.size or g[2][0] != o[2][0] or g[3][0] != o[3][0]:
            return t, flip, g[1][0][flip], o[1][0][flip]
    return None


def test_sum_product_tiles_bit_identical(shaped):
    code, llrs = shaped
    method = SumProduct(max_iterations=30)
    for esn0, x in llrs.items():
        want = gpu_decode(code, x, method, SP_TILES[0])
        for tiles in SP_TILES[1:] + [None]:
            got = gpu_decode(code, x, method, tiles)
            for name, g, w in zip(("bits", "llrs", "valid", "iterations"), got, want):
                bad = np.flatnonzero((g != w).reshape(g.shape[0], -1).any(axis=1))
                assert bad.size == 0, f"{esn0} dB {tiles}: {name} differ on frames {bad.tolist()}"


@pytest.mark.parametrize("tiles", [(2, 1, 1, 0), (2, 1, 2, 0)], ids=["P1", "P2-tuned"])
def test_sum_product_shaped_against_oracle(shaped, tiles):
    code, llrs = shaped
    method = SumProduct(max_iterations=30)
    report = []
    for esn0, x in llrs.items():
        got = gpu_decode(code, x, method, tiles)
        want = oracle_decode(code, x, method)
        probe = probe_decode(code, x, method)
        hard_differs = ((got[1] < 0) != (want[1] < 0)).any(axis=1)
        div = np.flatnonzero(hard_differs | (got[2] != want[2]) | (got[3] != want[3]))
        for f in div:
            tr = trace(code, x[f], method, tiles)
            assert tr is not None, f"{esn0} dB frame {f}: differs at 30 iterations, not before"
            t, flip, g_app, o_app = tr
            report.append(f"{esn0} dB frame {f}: iterations gpu {got[3][f]} oracle {want[3][f]}, "
                          f"first flip at iteration {t}, bits {flip[:8].tolist()}, "
                          f"APP gpu {g_app[:8].tolist()} oracle {o_app[:8].tolist()}")  # fmt: skip
            assert np.all(np.abs(g_app - o_app) <= NEAR_TIE), report[-1]
        # divergence is the exception: a handful of near ties at most
        assert div.size <= 4, f"{esn0} dB: {div.size} divergent frames"
        probe_div = ((probe[0] < 0) != (want[1] < 0)).any(axis=1) | (probe[1] != want[3])
        same = np.setdiff1d(np.arange(FRAMES), np.union1d(div, np.flatnonzero(probe_div)))
        g, o, pr = got[1][same], want[1][same], probe[0][same]
        small = np.abs(o) <= SMALL_APP
        np.testing.assert_allclose(g[small], o[small], rtol=RTOL, atol=ATOL,
                                   err_msg=f"{esn0} dB: APP of unsaturated bits")  # fmt: skip
        dg, dp = np.abs(g - o), np.abs(pr - o)
        qg, qp = np.quantile(dg, QUANTILES), np.quantile(dp, QUANTILES)
        print(f"sum-product {tiles} at {esn0} dB: {FRAMES - div.size} of {FRAMES} frames identical "
              f"in hard decisions and iterations (mean {got[3].mean():.2f}); |APP - oracle| "
              f"quantiles {QUANTILES}: GPU {np.round(qg, 5).tolist()}, phi-rounding probe "
              f"{np.round(qp, 5).tolist()}")  # fmt: skip
        assert np.all(qg <= CONDITIONING * qp + 1e-4), (esn0, qg.tolist(), qp.tolist())
    for line in report:
        print("divergent:", line)
