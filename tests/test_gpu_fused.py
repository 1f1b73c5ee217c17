"""Demapping fused into the decoder's frame load (GPU).

``ck_ldpc_decode_symbols`` must give exactly what the two-launch receive
tail gives: the closed-form max-log LLRs of the same symbols
(``compute_llr`` / ``llr_rows``), decoded by ``ck_ldpc_decode`` -- APP LLRs
bit for bit (signed zeros included), hard bits, iteration counts and valid
flags, for every tile of the quasi-cyclic layered and flooding kernels,
min-sum and sum-product, QPSK to 256-QAM, one noise variance or one per
frame.  The
two-launch tail itself is pinned to the oracle by the decoder and demapper
tests.
"""

import numpy as np
import pytest

from commkit_amd import Constellation, MinSum, SumProduct, compute_llr
from commkit_amd.coding import ldpc as ldpc_module
from commkit_amd.mapping.demap import llr_rows
from commkit_amd.workloads import seeded_code

from .codes import small_qc

pytestmark = pytest.mark.gpu


def _symbols(c, rows, per_row, snr_db, seed):
    """Noisy points of ``c``: ``(rows, per_row)`` complex64, with exact
    zeros, negative zeros and decision boundaries mixed in."""
    r = np.random.default_rng(seed)
    pts = c.points[r.integers(0, c.order, size=(rows, per_row))]
    nv = 10 ** (-snr_db / 10)
    x = pts + np.sqrt(nv / 2) * (r.standard_normal(pts.shape) + 1j * r.standard_normal(pts.shape))
    x = x.astype(np.complex64)
    flat = x.reshape(-1).view(np.float32).copy()
    hit = r.random(flat.size)
    flat[hit < 0.01] = 0.0
    flat[(hit >= 0.01) & (hit < 0.02)] = -0.0
    re_levels = np.unique(np.round(c.points.real, 6)).astype(np.float32)
    mids = (re_levels[1:] + re_levels[:-1]) / 2 if re_levels.size > 1 else re_levels
    pick = (hit >= 0.02) & (hit < 0.03)
    flat[pick] = mids[r.integers(0, mids.size, size=int(pick.sum()))]
    return flat.view(np.complex64).reshape(rows, per_row), nv


def _two_launch(code, c, x, nv, method, tiles):
    import torch

    if np.ndim(nv) == 0:
        llr = compute_llr(x, noise_var=float(nv), constellation=c, method="maxlog")
    else:
        llr = llr_rows(c, x, nv, "maxlog")
    llr = llr.reshape(x.shape[0], code.n).contiguous()
    out = ldpc_module._decode(code, llr, method, tiles=tiles)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out]


def _fused(code, c, x, nv, method, tiles):
    import torch

    out = ldpc_module._decode_symbols(code, x, c, nv, method, tiles=tiles)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out]


def _assert_same(got, want, label):
    for name, g, w in zip(("bits", "app", "valid", "iterations"), got, want):
        if g.dtype == np.float32:
            g, w = g.view(np.uint32), w.view(np.uint32)
        if not np.array_equal(g, w):
            bad = np.flatnonzero((g != w).reshape(g.shape[0], -1).any(axis=1))
            pytest.fail(f"{label}: {name} differ on frames {bad[:20].tolist()}")


# Quasi-cyclic layered tiles: (kernel, frames per group, lanes per row, state
# in LDS); (2, 0, 0, -1) lets the library choose.
QC_TILES = [(2, 0, 0, -1), (2, 1, 1, 0), (2, 1, 2, 0), (2, 1, 4, 0), (2, 1, 1, 1),
            (2, 2, 2, 1), (2, 3, 1, 0)]  # fmt: skip

CASES = [
    ("r4/5-z360", 64, 5, 17.0),
    ("r4/5-z360", 16, 3, 11.0),
    ("r1/2-z27", 4, 13, 2.0),
    ("r1/2-z27", 256, 9, 16.0),
    ("r2/3-z96", 64, 7, 13.0),
    ("r5/6-z360", 256, 3, 21.0),
]


@pytest.mark.parametrize("name,order,rows,snr_db", CASES)
def test_fused_equals_two_launches(gpu, name, order, rows, snr_db):
    import torch

    code = seeded_code(name)
    c = Constellation.qam(order)
    x_np, nv = _symbols(c, rows, code.n // c.bits_per_symbol, snr_db, seed=order + rows)
    x = torch.from_numpy(x_np).cuda()
    method = MinSum(scale=0.75, offset=0.0, max_iterations=12)
    for tiles in QC_TILES:
        try:
            want = _two_launch(code, c, x, nv, method, tiles)
        except RuntimeError:  # the tile does not fit this code: refused alike
            with pytest.raises(RuntimeError):
                _fused(code, c, x, nv, method, tiles)
            continue
        _assert_same(_fused(code, c, x, nv, method, tiles), want, f"{name} {order} {tiles}")


@pytest.mark.parametrize("name,order", [("r4/5-z360", 64), ("r1/2-z27", 16)])
def test_fused_sum_product(gpu, name, order):
    import torch

    code = seeded_code(name)
    c = Constellation.qam(order)
    x_np, nv = _symbols(c, 5, code.n // c.bits_per_symbol, 14.0 if order == 64 else 7.0, seed=3)
    x = torch.from_numpy(x_np).cuda()
    method = SumProduct(max_iterations=10)
    _assert_same(_fused(code, c, x, nv, method, None), _two_launch(code, c, x, nv, method, None),
                 f"{name} sum-product")  # fmt: skip


def test_fused_noise_per_frame(gpu):
    import torch

    code = seeded_code("r4/5-z360")
    c = Constellation.qam(64)
    rows = 7
    x_np, _ = _symbols(c, rows, code.n // 6, 16.0, seed=11)
    x = torch.from_numpy(x_np).cuda()
    nv = 10 ** (-np.linspace(14.0, 19.0, rows) / 10)
    nv[3] = 0.0  # clamped to 1e-20, as the demapper does
    method = MinSum(scale=0.75, max_iterations=15, early_stop=True)
    for tiles in (None, (2, 1, 1, 0), (2, 3, 1, 0)):
        _assert_same(_fused(code, c, x, nv, method, tiles),
                     _two_launch(code, c, x, nv, method, tiles), f"per-frame {tiles}")  # fmt: skip


def test_fused_small_code_qpsk(gpu):
    import torch

    code = small_qc()
    c = Constellation.qam(4)
    x_np, nv = _symbols(c, 9, code.n // 2, 3.0, seed=5)
    x = torch.from_numpy(x_np).cuda()
    method = MinSum(max_iterations=8, early_stop=True)
    _assert_same(_fused(code, c, x, nv, method, None), _two_launch(code, c, x, nv, method, None),
                 "small qpsk")  # fmt: skip


def test_fused_rejects(gpu):
    import torch

    code = seeded_code("r1/2-z27")
    c = Constellation.qam(64)
    x = torch.zeros((2, code.n // 6), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="quasi-cyclic kernel"):
        ldpc_module._decode_symbols(code, x, c, 0.1, MinSum(), tiles=(1, 0, 0, -1))
    with pytest.raises(RuntimeError, match="quasi-cyclic kernel"):
        ldpc_module._decode_symbols(code, x, c, 0.1, MinSum(schedule="flooding"),
                                    tiles=(1, 0, 0, -1))
    with pytest.raises(ValueError):
        ldpc_module._decode_symbols(code, x, Constellation.psk(8), 0.1, MinSum())
    with pytest.raises(ValueError):
        ldpc_module._decode_symbols(code, x, Constellation.qam(64).shaped(nu=0.05), 0.1, MinSum())
    with pytest.raises(ValueError):
        ldpc_module._decode_symbols(code, x[:, :-1], c, 0.1, MinSum())
    big = seeded_code("r4/5-z360")
    xb = torch.zeros((1, big.n // 10), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="1..4 bits"):
        ldpc_module._decode_symbols(big, xb, Constellation.qam(1024), 0.1, MinSum())
    bits, app, valid, iters = ldpc_module._decode_symbols(code, x[:0], c, 0.1, MinSum())
    assert bits.shape == (0, code.k) and app.shape == (0, code.n)


# Flooding tiles: (kernel, frames per group, lanes per row, state placement).
FLOOD_TILES = [(2, 0, 0, -1), (2, 1, 1, 0), (2, 1, 2, 0), (2, 2, 4, 0), (2, 3, 1, 0)]


@pytest.mark.parametrize("name,order,rows,snr_db", [("r1/2-z27", 16, 64, 8.0),
                                                    ("r1/2-z27", 256, 7, 16.0),
                                                    ("r2/3-z96", 64, 9, 13.0)])
def test_fused_flooding(gpu, name, order, rows, snr_db):
    # These codes run the flooding kernel's LDS form (the tables are staged
    # over the area its channel LLRs and states take after the load).
    import torch

    code = seeded_code(name)
    c = Constellation.qam(order)
    x_np, nv = _symbols(c, rows, code.n // c.bits_per_symbol, snr_db, seed=order + rows + 1)
    x = torch.from_numpy(x_np).cuda()
    for method in (MinSum(scale=0.75, max_iterations=20, schedule="flooding"),
                   SumProduct(max_iterations=8, schedule="flooding")):
        tiles_list = FLOOD_TILES if isinstance(method, MinSum) else [(2, 0, 0, -1), (2, 1, 1, 0)]
        for tiles in tiles_list:
            try:
                want = _two_launch(code, c, x, nv, method, tiles)
            except RuntimeError:  # the tile does not fit this code: refused alike
                with pytest.raises(RuntimeError):
                    _fused(code, c, x, nv, method, tiles)
                continue
            _assert_same(_fused(code, c, x, nv, method, tiles), want,
                         f"{name} {order} {type(method).__name__} {tiles}")


def test_fused_flooding_refuses_the_hbm_form(gpu):
    # r4/5-z360 (n = 18000) floods in the HBM form, which reads the channel
    # LLRs from the input every iteration: with symbols in, the call is
    # refused (checked on the host, before any launch).
    import torch

    code = seeded_code("r4/5-z360")
    x = torch.zeros((2, code.n // 6), dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="fit in LDS"):
        ldpc_module._decode_symbols(code, x, Constellation.qam(64), 0.1,
                                    MinSum(schedule="flooding"))  # fmt: skip


def test_fused_flooding_noise_per_frame(gpu):
    import torch

    code = seeded_code("r1/2-z27")
    c = Constellation.qam(16)
    rows = 11
    x_np, _ = _symbols(c, rows, code.n // 4, 8.0, seed=21)
    x = torch.from_numpy(x_np).cuda()
    nv = 10 ** (-np.linspace(6.0, 12.0, rows) / 10)
    method = MinSum(scale=0.75, max_iterations=20, schedule="flooding")
    _assert_same(_fused(code, c, x, nv, method, None), _two_launch(code, c, x, nv, method, None),
                 "flooding per-frame")  # fmt: skip
