"""The device front end: ck_ldpc_encode against the host encoder and the
parity checks, ck_awgn against the oracle's restated Philox stream and
Gaussian statistics, and generate / apply_awgn end to end."""

import math

import numpy as np
import pytest
from scipy import stats

from commkit_amd import Constellation, LDPCCode, compute_llr, decode
from commkit_amd.coding import CCDM, PAS
from commkit_amd.coding.ldpc import _encode
from commkit_amd.core import generate
from commkit_amd.impairments import apply_awgn
from commkit_amd.metrics import ber, fer
from commkit_amd.operating import waterfall_db
from commkit_amd.workloads import seeded_code
from oracle import cpu

from .codes import encode_host, small_qc, syndrome

pytestmark = pytest.mark.gpu


def _dev():
    import torch

    return torch.device("cuda", 0)


def _accumulate_code():
    m, k = 70, 90
    rng = np.random.default_rng(2)
    rows, cols = [], []
    for i in range(m):
        for j in rng.choice(k, 3, replace=False):
            rows.append(i)
            cols.append(j)
        rows.append(i)
        cols.append(k + i)
        if i:
            rows.append(i)
            cols.append(k + i - 1)
    return LDPCCode(np.array(rows), np.array(cols), m, m + k)


@pytest.mark.parametrize("which", ["small_qc", "accumulate", "r1/2-z27", "r5/6-z360",
                                   "punctured"])
@pytest.mark.parametrize("rows", [1, 17, 40])
def test_encoder_matches_the_host(which, rows):
    """Bit-exact against the host encoder; the codewords satisfy every check
    (all of H when nothing is punctured); last group of 16 partial."""
    import torch

    if which == "small_qc":
        code = small_qc()
    elif which == "accumulate":
        code = _accumulate_code()
    elif which == "punctured":
        base = seeded_code("r1/2-z27").__dict__["_qc"][0]
        code = LDPCCode.quasi_cyclic(base, lifting=27, punctured=27)
    else:
        code = seeded_code(which)
    assert code._encoder.kind in ("inverse", "accumulate")
    u = np.random.default_rng(rows).integers(0, 2, size=(rows, code.k), dtype=np.int8)
    got = _encode(code, torch.from_numpy(u).to(_dev())).cpu().numpy()
    want = encode_host(code, u)
    assert got.shape == (rows, code.n)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got[:, : code.k], u)
    if code.punctured == 0:
        assert not syndrome(code, got).any()


def test_encoder_refuses_host_rows():
    import torch

    code = small_qc()
    with pytest.raises(ValueError, match="device"):
        _encode(code, torch.zeros((1, code.k), dtype=torch.int8))


@pytest.mark.parametrize("complex_", [True, False])
def test_awgn_matches_the_restated_stream(complex_):
    """The noise the kernel adds is the oracle's Philox / Box-Muller stream
    (float32 log and sincos against float64: within 2e-5)."""
    import torch

    from commkit_amd import _native

    rows, length = 3, 1001  # ragged against the four normals of a draw
    dtype = torch.complex64 if complex_ else torch.float32
    x = torch.zeros((rows, length), dtype=dtype, device=_dev())
    std = torch.tensor([1.0, 0.5, 2.0], dtype=torch.float32, device=_dev())
    out = torch.empty_like(x)
    seed, stream = 0x0123456789ABCDEF, 5
    _native.check(_native.lib().ck_awgn(x.data_ptr(), rows, length, int(complex_),
                                        std.data_ptr(), seed, stream, out.data_ptr(),
                                        _native.stream_of(_dev())))  # fmt: skip
    per = 2 if complex_ else 1
    z = cpu.awgn_normals(seed, stream, rows * length * per).reshape(rows, length * per)
    got = torch.view_as_real(out).reshape(rows, -1) if complex_ else out
    want = z * np.array([1.0, 0.5, 2.0])[:, None]
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=2e-5)
    # in place, and a zero deviation leaves the samples as they were
    y = torch.randn((rows, length), device=_dev()).to(dtype)
    keep = y.clone()
    zero = torch.zeros(rows, dtype=torch.float32, device=_dev())
    _native.check(_native.lib().ck_awgn(y.data_ptr(), rows, length, int(complex_),
                                        zero.data_ptr(), seed, stream, y.data_ptr(),
                                        _native.stream_of(_dev())))  # fmt: skip
    assert torch.equal(y, keep)


def test_awgn_statistics():
    """2^21 complex samples: each component N(0, 1/2) by Kolmogorov-Smirnov,
    I and Q uncorrelated, a new seed a new draw."""
    import torch

    x = torch.zeros(1 << 21, dtype=torch.complex64, device=_dev())
    n = apply_awgn(x, noise_var=1.0, rng=3).cpu().numpy()
    for comp in (n.real, n.imag):
        assert stats.kstest(comp / math.sqrt(0.5), "norm").pvalue > 1e-3
    assert abs(np.mean(n.real * n.imag)) < 5 * 0.5 / math.sqrt(n.size)
    again = apply_awgn(x, noise_var=1.0, rng=3).cpu().numpy()
    np.testing.assert_array_equal(again, n)
    assert not np.array_equal(apply_awgn(x, noise_var=1.0, rng=4).cpu().numpy(), n)


def test_apply_awgn_levels():
    """Es/N0 from the measured symbol energy (complex N0, real N0/2), the
    SNR from the mean power, absolute per-channel variances; host arrays in,
    host arrays out."""
    rng = np.random.default_rng(0)
    qpsk = (rng.choice([-1.0, 1.0], (2, 1 << 18)) + 1j * rng.choice([-1.0, 1.0], (2, 1 << 18)))
    qpsk = (qpsk * np.array([[1.0], [2.0]]) / math.sqrt(2)).astype(np.complex64)  # Es 1 and 4
    out = apply_awgn(qpsk, esn0_db=10.0, sps=1, rng=1)
    assert isinstance(out, np.ndarray) and out.dtype == np.complex64
    var = np.mean(np.abs(out - qpsk) ** 2, axis=1)
    np.testing.assert_allclose(var, [0.1, 0.4], rtol=0.01)
    pam = rng.choice([-1.0, 1.0], 1 << 18).astype(np.float32)
    var = np.mean((apply_awgn(pam, esn0_db=3.0, sps=1, rng=2) - pam) ** 2)
    assert var == pytest.approx(10 ** -0.3 / 2, rel=0.01)
    var = np.mean(np.abs(apply_awgn(qpsk, snr_db=20.0, per_channel=False, rng=5) - qpsk) ** 2,
                  axis=1)  # fmt: skip
    np.testing.assert_allclose(var, [0.025, 0.025], rtol=0.01)
    var = np.mean(np.abs(apply_awgn(qpsk, noise_var=[0.01, 0.2], rng=6) - qpsk) ** 2, axis=1)
    np.testing.assert_allclose(var, [0.01, 0.2], rtol=0.01)
    with pytest.raises(ValueError, match="exactly one"):
        apply_awgn(qpsk, esn0_db=1.0, snr_db=1.0)
    with pytest.raises(ValueError, match="sps"):
        apply_awgn(qpsk, esn0_db=1.0)
    with pytest.raises(ValueError, match="one entry per"):
        apply_awgn(qpsk, noise_var=[1.0, 2.0, 3.0])


def test_generate_coded_record():
    """The reference holds the information bits, their codewords (the host
    encoder's), the symbols they map to; a seed repeats the record."""
    import torch

    code = seeded_code("r2/3-z96")
    c = Constellation.qam(16)
    sig = generate(c, symbol_rate=32e9, code=code, num_blocks=6, num_channels=2, rng=7,
                   device="gpu")  # fmt: skip
    ref = sig.reference
    assert sig.samples.is_cuda and sig.samples.shape == (2, 6 * code.n // 4)
    info = ref.info_bits.cpu().numpy()
    assert info.shape == (2, 6 * code.k)
    words = encode_host(code, info.reshape(-1, code.k)).reshape(2, -1)
    np.testing.assert_array_equal(ref.bits.cpu().numpy(), words)
    assert torch.equal(ref.symbols, c.map(ref.bits))
    host = generate(c, symbol_rate=32e9, code=code, num_blocks=6, num_channels=2, rng=7)
    assert isinstance(host.samples, np.ndarray)
    np.testing.assert_array_equal(host.reference.bits, words)
    padded = generate(c, symbol_rate=1e9, code=code, num_symbols=1000, pad=True, rng=1)
    assert padded.samples.shape == (1000,) and padded.reference.info_bits.size == 1 * code.k
    with pytest.raises(ValueError, match="pad=True"):
        generate(c, symbol_rate=1e9, code=code, num_symbols=1000, rng=1)
    with pytest.raises(ValueError, match="pulse"):
        generate(c, symbol_rate=1e9, num_symbols=8, sps=2)


def test_generate_shaped_record_has_its_composition():
    c = Constellation.qam(64).shaped(entropy=5.0)
    sig = generate(c, symbol_rate=1e9, num_symbols=4000, num_channels=2, rng=2)
    code = sig.reference.code
    assert isinstance(code, CCDM) and code.num_symbols == 4000
    labels = sig.reference.bits.reshape(2, 4000, 6) @ (1 << np.arange(5, -1, -1))
    pts_by_label = np.empty(64, dtype=np.int64)
    pts_by_label[np.asarray(c.bit_labels) @ (1 << np.arange(5, -1, -1))] = np.arange(64)
    for row in labels:
        np.testing.assert_array_equal(np.bincount(pts_by_label[row], minlength=64),
                                      code.composition)  # fmt: skip
    with pytest.raises(ValueError, match="PAS"):
        generate(c, symbol_rate=1e9, code=seeded_code("r2/3-z96"), num_blocks=3)


@pytest.mark.parametrize("kind", ["ldpc", "pas"])
def test_end_to_end(kind):
    """generate -> apply_awgn -> compute_llr -> decode -> ber/fer, every step
    on the device, above the code's waterfall: error free."""
    if kind == "ldpc":
        c, code = Constellation.qam(64), seeded_code("r4/5-z360")
        es = waterfall_db("r4/5-z360", 64) + 1.0
        method = None
    else:
        from commkit_amd import SumProduct

        c = Constellation.qam(256).shaped(nu=0.05)
        code, es, method = PAS(seeded_code("r5/6-z360"), c), 14.5, SumProduct(max_iterations=30)
    sig = generate(c, symbol_rate=64e9, code=code, num_blocks=8, rng=11, device="gpu")
    rx = apply_awgn(sig, esn0_db=es, rng=12)
    llrs = compute_llr(rx, esn0_db=es, method="exact" if kind == "pas" else "maxlog")
    res = decode(llrs, code=sig.reference.code, method=method)
    assert fer(res, rx) == 0.0 and ber(res, rx) == 0.0
    assert ber(rx) > 0  # the channel did flip bits before decoding
