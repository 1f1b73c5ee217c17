"""HIP LDPC decoder against the CPU oracle (GPU).

Min-sum (layered and flooding) must equal the oracle bit for bit: APP LLRs,
hard bits, iteration counts and valid flags, on every frame and for every
kernel and tile choice.  Sum-product matches to the reference's own GPU/CPU
tolerance (rtol 1e-3, atol 1e-3, ``tests/coding/test_ldpc_kernel.py:37-45``):
``phi`` uses ocml's logf/tanhf on the GPU and glibc's on the CPU; frames whose
hard bits or iteration counts differ are listed.
"""

import numpy as np
import pytest

from commkit_amd import LDPCCode, MinSum, SumProduct, decode
from commkit_amd.coding import ldpc as ldpc_module
from commkit_amd.workloads import seeded_code
from oracle import cpu

from .codes import encode_host, quiet, small_qc

pytestmark = pytest.mark.gpu


def _llrs(code, rows, mean, seed):
    r = np.random.default_rng(seed)
    return (mean + 1.5 * r.standard_normal((rows, code.n))).astype(np.float32)


def gpu_decode(code, llrs, method, tiles=None):
    import torch

    x = torch.from_numpy(llrs).cuda()
    bits, app, valid, iters = ldpc_module._decode(code, x, method, tiles=tiles)
    torch.cuda.synchronize()
    return bits.cpu().numpy(), app.cpu().numpy(), valid.cpu().numpy(), iters.cpu().numpy()


def oracle_decode(code, llrs, method):
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


def assert_identical(got, want, label=""):
    names = ("bits", "llrs", "valid", "iterations")
    for name, g, w in zip(names, got, want):
        if not np.array_equal(g, w):
            bad = np.flatnonzero((g != w).reshape(g.shape[0], -1).any(axis=1))
            pytest.fail(f"{label}: {name} differ on frames {bad[:20].tolist()}")


CODES = {
    "small": (small_qc, 7),
    "r1/2-z27": (lambda: seeded_code("r1/2-z27"), 64),
    "r2/3-z96": (lambda: seeded_code("r2/3-z96"), 16),
    "r4/5-z360": (lambda: seeded_code("r4/5-z360"), 6),
    "r5/6-z360": (lambda: seeded_code("r5/6-z360"), 4),
}


@pytest.mark.parametrize("name", list(CODES))
@pytest.mark.parametrize("early_stop", [False, True])
def test_min_sum_layered_bit_exact(gpu, name, early_stop):
    make, rows = CODES[name]
    code = make()
    llrs = _llrs(code, rows, 1.8, seed=1)
    method = MinSum(scale=0.75, offset=0.05, max_iterations=12, early_stop=early_stop)
    assert_identical(gpu_decode(code, llrs, method), oracle_decode(code, llrs, method), name)


# (kernel, frames per group, lanes per row, state in LDS): generic CSR, and
# the QC kernel across tiles.
TILES = [(1, 0, 0, -1), (2, 1, 1, 0), (2, 1, 2, 0), (2, 1, 4, 0), (2, 1, 1, 1), (2, 2, 2, 1),
         (2, 3, 1, 0), (2, 0, 0, -1)]  # fmt: skip


@pytest.mark.parametrize("name", ["r1/2-z27", "r2/3-z96", "r4/5-z360"])
def test_tiles_bit_identical(gpu, name):
    make, rows = CODES[name]
    code = make()
    llrs = _llrs(code, min(rows, 12), 1.6, seed=2)
    method = MinSum(max_iterations=10)
    want = oracle_decode(code, llrs, method)
    for tiles in TILES:
        if name == "r4/5-z360" and tiles[0] == 2 and tiles[1] * 18000 * 4 > 160 * 1024:
            continue
        assert_identical(gpu_decode(code, llrs, method, tiles), want, f"{name} {tiles}")


@pytest.mark.parametrize("name", ["small", "r1/2-z27", "r2/3-z96", "r4/5-z360"])
@pytest.mark.parametrize("early_stop", [False, True])
def test_min_sum_flooding_bit_exact(gpu, name, early_stop):
    # The three small codes keep their check states, channel LLRs and column
    # tables in LDS; r4/5-z360 (n = 18000) keeps them in HBM.
    make, rows = CODES[name]
    code = make()
    llrs = _llrs(code, rows, 1.8, seed=3)
    method = MinSum(scale=0.75, max_iterations=20, early_stop=early_stop, schedule="flooding")
    want = oracle_decode(code, llrs, method)
    for tiles in (None, (1, 0, 0, -1), (2, 1, 2, -1)):
        assert_identical(gpu_decode(code, llrs, method, tiles), want, f"{name} {tiles}")


@pytest.mark.parametrize("schedule", ["layered", "flooding"])
@pytest.mark.parametrize("name", ["r1/2-z27", "r4/5-z360"])
def test_sum_product_matches(gpu, name, schedule):
    make, _ = CODES[name]
    code = make()
    llrs = _llrs(code, 4, 0.8, seed=4)
    method = SumProduct(max_iterations=5, early_stop=False, schedule=schedule)
    want = oracle_decode(code, llrs, method)
    for tiles in (None, (1, 0, 0, -1)):
        got = gpu_decode(code, llrs, method, tiles)
        np.testing.assert_allclose(got[1], want[1], rtol=1e-3, atol=1e-3)
        diverged = np.flatnonzero((got[0] != want[0]).any(axis=1) | (got[3] != want[3]))
        print(f"sum-product {name} {schedule} {tiles}: divergent frames {diverged.tolist()}")


@pytest.mark.parametrize("name,lo,hi,rows", [("r1/2-z27", 1.9, 2.6, 16), ("r4/5-z360", 3.0, 3.8, 8)])
def test_sum_product_flooding_early_stop(gpu, name, lo, hi, rows):
    # The flooding kernel's stop rule is the parity of the hard decisions its
    # check phase reads.  Frames with mean LLRs across each code's waterfall
    # converge after 2 to 16 iterations or not at all (the oracle runs the
    # first four r1/2-z27 frames and the first r4/5-z360 frame to the limit);
    # each must stop where the oracle stops, with its bits and valid flag.
    code = CODES[name][0]()
    r = np.random.default_rng(14)
    mean = np.linspace(lo, hi, rows)[:, None]
    llrs = (mean + 1.5 * r.standard_normal((rows, code.n))).astype(np.float32)
    method = SumProduct(max_iterations=25, early_stop=True, schedule="flooding")
    want = oracle_decode(code, llrs, method)
    got = gpu_decode(code, llrs, method)
    assert len(np.unique(want[3])) > 1, "the frames should stop at different iterations"
    np.testing.assert_array_equal(got[3], want[3])
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_allclose(got[1], want[1], rtol=1e-3, atol=1e-3)


def test_noiseless_and_random(gpu):
    code = seeded_code("r2/3-z96")
    u = np.random.default_rng(5).integers(0, 2, (5, code.k))
    words = encode_host(code, u)
    clean = ((1.0 - 2.0 * words) * 4.0).astype(np.float32)
    for method in (MinSum(max_iterations=4), SumProduct(max_iterations=4),
                   MinSum(max_iterations=4, schedule="flooding")):  # fmt: skip
        bits, app, valid, iters = gpu_decode(code, clean, method)
        assert valid.all() and (iters == 0).all()
        np.testing.assert_array_equal(bits, u)
        np.testing.assert_array_equal(app, clean)
        noise = np.random.default_rng(6).standard_normal((5, code.n)).astype(np.float32)
        _, _, valid, iters = gpu_decode(code, noise, method)
        assert not valid.any() and (iters == 4).all()


def test_frame_alone_equals_frame_in_batch(gpu):
    code = seeded_code("r1/2-z27")
    llrs = _llrs(code, 40, 1.2, seed=7)
    method = MinSum(max_iterations=20)
    for tiles in (None, (2, 9, 1, 0), (2, 4, 2, 1), (1, 0, 0, -1)):
        batch = gpu_decode(code, llrs, method, tiles)
        for f in (0, 13, 39):
            alone = gpu_decode(code, llrs[f : f + 1], method, tiles)
            for g, w in zip(alone, batch):
                np.testing.assert_array_equal(g[0], w[f])


def test_chunking_agrees(gpu, monkeypatch):
    code = seeded_code("r2/3-z96")
    llrs = _llrs(code, 9, 1.5, seed=8)
    method = SumProduct(max_iterations=6, early_stop=False)
    whole = gpu_decode(code, llrs, method)
    monkeypatch.setattr(ldpc_module, "_GPU_WORKSPACE", code.checks.size * 4 * 2)
    part = gpu_decode(code, llrs, method)
    for g, w in zip(part, whole):
        np.testing.assert_array_equal(g, w)


def test_punctured_bits_are_erasures(gpu):
    code = quiet(LDPCCode.quasi_cyclic, seeded_code("r2/3-z96").__dict__["_qc"][0],
                 lifting=96, punctured=96)  # fmt: skip
    assert code.n == 2304 - 96
    u = np.random.default_rng(9).integers(0, 2, (3, code.k))
    words = encode_host(code, u)
    llrs = ((1.0 - 2.0 * words) * 2.0).astype(np.float32)
    res = decode(llrs.reshape(-1), code=code, method=MinSum(max_iterations=20))
    np.testing.assert_array_equal(res.bits, u.reshape(-1))
    assert res.valid.all() and res.llrs.shape == (3 * code.n,)


@pytest.mark.parametrize("schedule", ["layered", "flooding"])
def test_odd_length_code_bit_exact(gpu, schedule):
    # n = 35 (the small QC code with its last 5 variables punctured): the
    # frame store takes its one-float form instead of 16-byte vectors, and
    # the loads their scalar form; every kernel and tile still equals the
    # oracle bit for bit.
    code = quiet(LDPCCode.quasi_cyclic, small_qc().__dict__["_qc"][0], lifting=5, punctured=5)
    assert code.n % 4 == 3
    llrs = _llrs(code, 9, 1.6, seed=15)
    method = MinSum(scale=0.75, max_iterations=12, schedule=schedule)
    want = oracle_decode(code, llrs, method)
    for tiles in (None, (1, 0, 0, -1), (2, 1, 1, 0), (2, 2, 2, 0), (2, 3, 4, 0)):
        assert_identical(gpu_decode(code, llrs, method, tiles), want, f"{schedule} {tiles}")


def test_decode_contract(gpu):
    import torch

    code = seeded_code("r1/2-z27")
    llrs = _llrs(code, 3, 2.0, seed=10)
    host = decode(llrs.reshape(-1), code=code)
    assert isinstance(host.bits, np.ndarray) and host.bits.dtype == np.int8
    assert host.bits.shape == (3 * code.k,) and host.llrs.shape == (3 * code.n,)
    assert host.valid.dtype == bool and host.iterations.dtype == np.int32
    assert host.valid.shape == (3,) and host.checked
    dev = decode(torch.from_numpy(llrs).cuda(), code=code)
    assert dev.bits.device.type == "cuda" and dev.bits.shape == (3, code.k)
    np.testing.assert_array_equal(dev.bits.cpu().numpy().reshape(-1), host.bits)
    two = decode(llrs.reshape(1, -1), code=code)
    assert two.bits.shape == (1, 3 * code.k) and two.valid.shape == (1, 3)
    hard = decode((llrs < 0).astype(np.int8).reshape(-1), code=code,
                  method=MinSum(max_iterations=3))  # fmt: skip
    assert hard.llrs.dtype == np.float32
    empty = decode(np.zeros((2, 0), np.float32), code=code)
    assert empty.bits.shape == (2, 0) and empty.valid.shape == (2, 0)
    with pytest.raises(ValueError, match="multiple of the block length"):
        decode(np.zeros(code.n + 1, np.float32), code=code)
    from commkit_amd import PrecisionWarning

    with pytest.warns(PrecisionWarning):
        decode(llrs.astype(np.float64).reshape(-1), code=code)


def test_shaped_pas_frames_decode(gpu):
    """Probabilistic amplitude shaping frames end to end at a high Es/N0:
    exact LLRs with the prior, gathered into codeword order, decode by
    layered sum-product to the codeword's information part (the matcher's
    amplitude bits, then the payload's sign bits)."""
    import torch

    from commkit_amd import Constellation, SumProduct
    from commkit_amd.mapping.demap import llr_any
    from commkit_amd.workloads import make_pas_frames, seeded_code

    code = seeded_code("r5/6-z360")
    c = Constellation.qam(256).shaped(nu=0.05)
    fr = make_pas_frames(code, c, 16.0, 6, seed=3)
    assert abs(float((fr.tx.abs() ** 2).mean()) - c.power()) < 0.02
    llr = llr_any(c, fr.rx, float(fr.noise_var[0]), "exact").reshape(6, code.n)
    llr_c = llr[:, fr.iperm].contiguous()
    bits, app, valid, iters = gpu_decode(code, llr_c.cpu().numpy(), SumProduct(max_iterations=30))
    assert valid.all()
    words = fr.sent[:, fr.iperm]  # codeword order
    np.testing.assert_array_equal(bits, words[:, : code.k].cpu().numpy())  # systematic
    ccdm = fr.pas.matcher
    assert torch.equal(words[:, ccdm.n : code.k], fr.info[:, ccdm.k :])  # payload on signs


def test_infeasible_tuned_tile_falls_back(gpu, monkeypatch):
    """A tuned tile the library cannot run for a code (here the check state
    in LDS beside two 72 KB APPs) is a hint: decode takes the default tile
    and gives the same results; the same tile asked for explicitly fails."""
    code = seeded_code("r4/5-z360")
    key = (ldpc_module.code_fingerprint(code), 0, 0)
    monkeypatch.setitem(ldpc_module.TUNED, key, (2, 2, 1, 1))
    llrs = _llrs(code, 3, 1.6, seed=11)
    method = MinSum(max_iterations=8)
    assert_identical(gpu_decode(code, llrs, method), gpu_decode(code, llrs, method, (0, 0, 0, -1)),
                     "fallback")  # fmt: skip
    assert_identical(gpu_decode(code, llrs, method), oracle_decode(code, llrs, method), "oracle")
    with pytest.raises(RuntimeError, match="LDS"):
        gpu_decode(code, llrs, method, (2, 2, 1, 1))


@pytest.mark.parametrize("schedule", ["layered", "flooding"])
def test_signed_zero_llrs(gpu, schedule):
    """Channel LLRs of exactly -0.0 and +0.0 (a symbol on a decision
    boundary): the decoder takes -0.0 as +0.0, as every comparison of the
    reference does, so bits, iterations and valid equal the oracle's and the
    APP LLRs equal it as values (a zero APP may differ in its sign bit)."""
    code = seeded_code("r2/3-z96")
    llrs = _llrs(code, 6, 1.2, seed=12)
    r = np.random.default_rng(13)
    hit = r.random(llrs.shape) < 0.05
    llrs[hit] = np.where(r.random(hit.sum()) < 0.5, -0.0, 0.0).astype(np.float32)
    assert np.signbit(llrs[hit]).any() and (~np.signbit(llrs[hit])).any()
    method = MinSum(scale=0.75, max_iterations=15, schedule=schedule)
    want = oracle_decode(code, llrs, method)
    for tiles in (None, (2, 1, 1, 0), (2, 1, 2, 0), (2, 1, 4, 0)):
        got = gpu_decode(code, llrs, method, tiles)
        for name, g, w in zip(("bits", "llrs", "valid", "iterations"), got, want):
            assert np.array_equal(g, w), f"{schedule} {tiles}: {name}"
