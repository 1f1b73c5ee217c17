"""Scoring on the GPU: ck_bit_errors and ck_gmi_loss against float64 NumPy
counts and sums, closed forms (error-free and zero LLRs), the binary-input
AWGN mutual information by quadrature, and archives loaded onto the GPU."""

import numpy as np
import pytest

from commkit_amd import Constellation, Reference, Signal, compute_llr, decode
from commkit_amd.io import load_npz, save_npz
from commkit_amd.metrics import ber, fer, gmi, ngmi
from commkit_amd.workloads import make_frames, seeded_code

pytestmark = pytest.mark.gpu

PS256 = Constellation.qam(256).shaped(nu=0.05)


def _dev():
    import torch

    return torch.device("cuda", 0)


def _loss64(llrs, bits):
    """The mean of log2(1 + e^{-(1 - 2c) L}) per row in float64."""
    x = -np.asarray(llrs, np.float64) * (1.0 - 2.0 * np.asarray(bits, np.float64))
    return np.mean(np.logaddexp(0.0, x), axis=-1) / np.log(2.0)


@pytest.mark.parametrize(("rows", "n", "block"), [(1, 1000, 250), (3, 36000, 18000),
                                                  (2, 4 * 1037, 1037), (5, 16, 16)])
def test_bit_and_block_errors(rows, n, block):
    """Ragged lengths (no 16-byte multiple), blocks with and without errors."""
    import torch

    rng = np.random.default_rng(rows * n)
    tx = rng.integers(0, 2, size=(rows, n), dtype=np.int8)
    flips = rng.random((rows, n)) < 0.002
    flips[:, :block] = False  # the first block of every row is clean
    rx = tx ^ flips.astype(np.int8)
    want_ber = flips.sum(axis=1) / n
    want_fer = flips.reshape(rows, -1, block).any(axis=2).mean(axis=1)
    got_ber = ber(torch.from_numpy(rx).to(_dev()), tx)
    got_fer = fer(rx, torch.from_numpy(tx).to(_dev()), info_bits=block)
    assert isinstance(got_ber, np.ndarray) and got_ber.shape == (rows,)
    np.testing.assert_array_equal(got_ber, want_ber)
    np.testing.assert_array_equal(got_fer, want_fer)
    one = ber(rx[0], tx[0])
    assert isinstance(one, float) and one == want_ber[0]
    # bool and unaligned (skipped) input: the byte loop instead of 16-byte loads
    c = Constellation.qam(16)
    got = ber(rx.astype(bool), tx, constellation=c, num_skip_symbols=3)
    np.testing.assert_array_equal(got, flips[:, 12:].sum(axis=1) / (n - 12))


def test_post_fec_rates_of_a_decode():
    """fer/ber of a DecodeResult against a Reference holding the code: k from
    the code, rates equal to counting the decoded frames by hand."""
    code = seeded_code("r2/3-z96")
    c = Constellation.qam(16)
    fr = make_frames(code, c, esn0_db=np.linspace(5.0, 11.0, 40), num_frames=40, seed=4)
    llrs = compute_llr(fr.rx, noise_var=float(fr.noise_var.mean()), constellation=c)
    res = decode(llrs.reshape(-1), code=code)
    info = fr.info.reshape(-1)
    sig = Signal(fr.rx.reshape(-1), constellation=c,
                 reference=Reference(symbols=fr.tx.reshape(-1), bits=fr.code_bits.reshape(-1),
                                     info_bits=info, code=code))  # fmt: skip
    wrong = (res.bits.reshape(40, code.k) != fr.info).cpu().numpy()
    assert 0 < wrong.any(axis=1).sum() < 40
    assert fer(res, sig) == wrong.any(axis=1).mean()
    assert ber(res, sig.reference) == wrong.mean()
    assert ber(res, info) == wrong.mean()
    with pytest.raises(ValueError, match="disagrees"):
        fer(res, sig, info_bits=code.k + 1)


def test_pre_fec_ber_of_a_signal():
    """A Signal is decided at the nearest point (the prior of a shaped
    constellation plays no part) and its labels compared with the bits."""
    import torch

    for c, es in ((Constellation.qam(64), 16.0), (PS256, 20.0)):
        rng = np.random.default_rng(1)
        bits = rng.integers(0, 2, size=(2, 6000 * c.bits_per_symbol), dtype=np.int8)
        tx = c.map(torch.from_numpy(bits).to(_dev())).cpu().numpy()
        nv = c.power() * 10 ** (-es / 10)
        rx = (tx + np.sqrt(nv / 2) * (rng.standard_normal(tx.shape)
                                      + 1j * rng.standard_normal(tx.shape))).astype(np.complex64)  # fmt: skip
        pts = np.asarray(c.points)
        nearest = np.abs(rx[..., None] - pts).argmin(axis=-1)
        decided = np.asarray(c.bit_labels)[nearest].reshape(2, -1)
        want = (decided != bits).mean(axis=1)
        assert want.min() > 0
        sig = Signal(rx, constellation=c, reference=Reference(bits=bits))
        np.testing.assert_array_equal(ber(sig), want)
        np.testing.assert_array_equal(
            ber(sig, num_skip_symbols=100),
            (decided[:, 100 * c.bits_per_symbol:] != bits[:, 100 * c.bits_per_symbol:]).mean(
                axis=1))


def test_gmi_closed_forms():
    """Error-free LLRs carry all of H(X); zero LLRs carry nothing
    (loss exactly 1 bit per bit)."""
    import torch

    rng = np.random.default_rng(2)
    for c in (Constellation.qam(16), PS256):
        k = c.bits_per_symbol
        bits = rng.integers(0, 2, size=(2, 900 * k), dtype=np.int8)
        sure = torch.from_numpy(60.0 * (1.0 - 2.0 * bits).astype(np.float32)).to(_dev())
        np.testing.assert_array_equal(ngmi(sure, bits, constellation=c), [1.0, 1.0])
        np.testing.assert_allclose(gmi(sure, bits, constellation=c), c.entropy, rtol=0,
                                   atol=1e-12)  # fmt: skip
        zero = torch.zeros(bits.shape, dtype=torch.float32, device=_dev())
        # log1p(1) / ln 2 in float32: one within an ulp of 1
        np.testing.assert_allclose(ngmi(zero, bits, constellation=c), [0.0, 0.0], atol=1.2e-7)
        np.testing.assert_allclose(gmi(zero, bits, constellation=c), c.entropy - k,
                                   atol=k * 1.2e-7)  # fmt: skip


@pytest.mark.parametrize(("rows", "n"), [(1, 6 * 999), (3, 6 * 40_001)])
def test_gmi_loss_matches_float64(rows, n):
    """Several workgroups a row (spans of 65,536 LLRs) and a ragged tail."""
    import torch

    rng = np.random.default_rng(n)
    bits = rng.integers(0, 2, size=(rows, n), dtype=np.int8)
    mu = 3.0
    llrs = ((1.0 - 2.0 * bits) * mu + np.sqrt(2 * mu) * rng.standard_normal((rows, n)))
    llrs = llrs.astype(np.float32)
    want = 1.0 - _loss64(llrs, bits)
    got = ngmi(torch.from_numpy(llrs).to(_dev()), bits, constellation=Constellation.qam(64))
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-7)
    skip = ngmi(llrs, bits, constellation=Constellation.qam(64), num_skip_symbols=7)
    np.testing.assert_allclose(skip, 1.0 - _loss64(llrs[:, 42:], bits[:, 42:]), rtol=0, atol=2e-7)
    if rows == 1:
        assert isinstance(ngmi(llrs[0], bits[0], constellation=Constellation.qam(64)), float)


def test_gmi_of_bpsk_is_the_biawgn_capacity():
    """BPSK with exact LLRs: GMI = 1 - E[log2(1 + e^{-2y/s2})] for y = 1 + n,
    by quadrature; 2^21 symbols give about 1e-3 of Monte Carlo spread."""
    import torch
    from scipy import integrate

    es_db = 0.0
    c = Constellation.pam(2)
    s2 = c.power() / 10 ** (es_db / 10) / 2  # real symbols: the in-phase N0 / 2

    def integrand(y):
        p = np.exp(-((y - 1.0) ** 2) / (2 * s2)) / np.sqrt(2 * np.pi * s2)
        return p * np.logaddexp(0.0, -2.0 * y / s2) / np.log(2.0)

    capacity = 1.0 - integrate.quad(integrand, -30, 30, limit=200)[0]
    g = torch.Generator(device=_dev())
    g.manual_seed(11)
    nsym = 1 << 21
    bits = torch.randint(0, 2, (nsym,), generator=g, device=_dev(), dtype=torch.int8)
    tx = c.map(bits)
    rx = (tx + np.sqrt(s2) * torch.randn(nsym, generator=g, device=_dev())).to(torch.float32)
    sig = Signal(rx, constellation=c, reference=Reference(symbols=tx, bits=bits))
    got = gmi(sig, esn0_db=es_db, method="exact")
    assert abs(got - capacity) < 3e-3
    assert got == pytest.approx(gmi(sig, esn0_db=es_db), abs=1e-6)  # max-log is exact for 2 points


def test_signal_is_scaled_by_its_data_aided_gain():
    """gmi of a Signal received at half amplitude equals gmi of its LLRs
    computed on samples divided by |<r s*>| / <|s|^2>."""
    import torch

    c = Constellation.qam(16)
    fr = make_frames(seeded_code("r2/3-z96"), c, esn0_db=10.0, num_frames=8, seed=3)
    rx = fr.rx * 0.5
    sig = Signal(rx, constellation=c, reference=Reference(symbols=fr.tx, bits=fr.code_bits))
    r, s = rx.to(torch.complex128), fr.tx.to(torch.complex128)
    gain = ((r * s.conj()).mean(dim=1, keepdim=True).abs() / (s.abs() ** 2).mean(dim=1,
                                                                                 keepdim=True))  # fmt: skip
    llrs = compute_llr((rx / gain.to(torch.float32)).to(torch.complex64), noise_var=0.1,
                       constellation=c)  # fmt: skip
    want = gmi(llrs, fr.code_bits, constellation=c)
    got = gmi(sig, noise_var=0.1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)
    assert np.all(got > 3.0) and np.all(got <= 4.0)


def test_archive_loads_onto_the_gpu(tmp_path):
    code = seeded_code("r2/3-z96")
    c = Constellation.qam(16)
    fr = make_frames(code, c, esn0_db=11.0, num_frames=4, seed=5)
    sig = Signal(fr.rx, constellation=c,
                 reference=Reference(symbols=fr.tx, bits=fr.code_bits, info_bits=fr.info,
                                     code=code))  # fmt: skip
    save_npz(sig, tmp_path / "cap.npz")
    back = load_npz(tmp_path / "cap.npz", device="gpu")
    assert back.samples.is_cuda and back.reference.info_bits.is_cuda
    assert back.reference.code.__dict__["_qc"][1] == 96
    np.testing.assert_array_equal(ber(back), ber(sig))
    llrs = compute_llr(back, noise_var=float(fr.noise_var[0]))
    res = decode(llrs, code=back.reference.code)
    np.testing.assert_array_equal(fer(res, back), np.zeros(4))


def test_ber_of_a_long_row():
    """One row of 5,000,011 bits: ber splits it over spans of 65,536 bits
    (the last one short) and still counts every error."""
    import torch

    rng = np.random.default_rng(8)
    tx = rng.integers(0, 2, size=5_000_011, dtype=np.int8)
    flips = rng.random(tx.size) < 1e-3
    rx = tx ^ flips.astype(np.int8)
    assert ber(torch.from_numpy(rx).to(_dev()), tx) == flips.sum() / tx.size
    assert ber(rx[3:], tx[3:]) == flips[3:].sum() / (tx.size - 3)  # unaligned start
