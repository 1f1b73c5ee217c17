"""Operating-point machinery on the GPU: awgn_gmi (Gauss-Hermite nodes
through the exact demapper kernel) against a float64 restatement of the
quadrature and against Monte Carlo, Monte-Carlo error-rate curves, and the
EXIT-predicted Es/N0 against the measured waterfalls."""

import math

import numpy as np
import pytest
from scipy.special import erfc

from commkit_amd import Constellation, Reference, Signal
from commkit_amd.coding import PAS
from commkit_amd.coding.performance import error_rate_curve
from commkit_amd.metrics import awgn_gmi, gmi
from commkit_amd.operating import threshold_esn0_db, waterfall_db
from commkit_amd.workloads import seeded_code

pytestmark = pytest.mark.gpu


def _gmi64(c, esn0_db, nodes=128):
    """GMI by the same Gauss-Hermite rule, every term in float64 on the host:
    H - sum_b sum_i p_i sum_q W_q log2(sum_j p_j e_ij / sum_{j: c_jb = c_ib}
    p_j e_ij), e_ij = exp(-(|x_i - x_j + n_q|^2 - |n_q|^2) / N0)."""
    x = np.asarray(c.points, np.complex128)
    p = np.full(c.order, 1 / c.order) if c.pmf is None else np.asarray(c.pmf)
    lab = np.asarray(c.bit_labels)
    n0 = c.power() / 10 ** (esn0_db / 10)
    t, w = np.polynomial.hermite.hermgauss(nodes)
    if c.is_complex:
        nq = (math.sqrt(n0) * (t[:, None] + 1j * t[None, :])).ravel()
        wq = (w[:, None] * w[None, :]).ravel() / math.pi
    else:
        nq, wq = math.sqrt(n0) * t + 0j, w / math.sqrt(math.pi)
    loss = 0.0
    for i in range(c.order):
        d = x[i] - x[None, :] + nq[:, None]  # (Q, M)
        lg = -(np.abs(d) ** 2 - np.abs(nq[:, None]) ** 2) / n0 + np.log(p)[None, :]
        top = np.logaddexp.reduce(lg, axis=1)
        for b in range(c.bits_per_symbol):
            same = lab[:, b] == lab[i, b]
            loss += p[i] * np.dot(wq, top - np.logaddexp.reduce(lg[:, same], axis=1))
    return c.entropy - loss / math.log(2)


@pytest.mark.parametrize(("c", "esn0"), [
    (Constellation.qam(16), [-5.0, 5.0, 15.0]),
    (Constellation.psk(8), [0.0, 10.0]),
    (Constellation.qam(64).shaped(entropy=5.0), [12.0]),
    (Constellation.pam(4, unipolar=True), [5.0]),
], ids=["qam16", "psk8", "qam64-shaped", "pam4-unipolar"])
def test_awgn_gmi_matches_float64_quadrature(c, esn0):
    got = awgn_gmi(c, esn0_db=esn0)
    want = [_gmi64(c, e) for e in esn0]
    assert isinstance(got, np.ndarray) and got.shape == (len(esn0),)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    assert isinstance(awgn_gmi(c, esn0_db=esn0[0]), float)


@pytest.mark.parametrize("c", [Constellation.qam(16), Constellation.qam(64).shaped(entropy=5.0)],
                         ids=["qam16", "qam64-shaped"])
def test_awgn_gmi_limits(c):
    """At high SNR the GMI is H(X); at -30 dB it tends to H(X) - sum_b H(B_b)
    (0 for a uniform prior, negative for a shaped one)."""
    g = awgn_gmi(c, esn0_db=np.array([-30.0, 5.0, 40.0]))
    assert np.all(g <= c.entropy + 1e-9)
    assert g[-1] == pytest.approx(c.entropy, abs=1e-5)
    pmf = np.full(c.order, 1 / c.order) if c.pmf is None else np.asarray(c.pmf)
    p1 = pmf @ np.asarray(c.bit_labels)
    bits_h = -np.sum(p1 * np.log2(p1) + (1 - p1) * np.log2(1 - p1))
    assert g[0] == pytest.approx(c.entropy - bits_h, abs=0.01)


def test_awgn_gmi_against_monte_carlo():
    """The measured gmi of 2^18 shaped 64-QAM symbols at 16 dB, exact LLRs,
    within 0.01 bit of the quadrature."""
    import torch

    c = Constellation.qam(64).shaped(entropy=5.0)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(21)
    pmf = torch.from_numpy(np.asarray(c.pmf)).to(dev)
    idx = torch.multinomial(pmf, 1 << 18, replacement=True, generator=g)
    bits = torch.from_numpy(np.asarray(c.bit_labels, np.int8)).to(dev)[idx].reshape(-1)
    tx = c.map(bits)
    n0 = c.power() / 10 ** 1.6
    noise = torch.complex(torch.randn(tx.shape, generator=g, device=dev),
                          torch.randn(tx.shape, generator=g, device=dev))  # fmt: skip
    rx = (tx + noise * math.sqrt(n0 / 2)).to(torch.complex64)
    measured = gmi(Signal(rx, constellation=c, reference=Reference(symbols=tx, bits=bits)),
                   esn0_db=16.0, method="exact")  # fmt: skip
    assert abs(measured - awgn_gmi(c, esn0_db=16.0)) < 0.01


def test_error_rate_curve():
    """QPSK on the rate-1/2 code around its waterfall: pre-FEC BER of Gray
    QPSK is Q(sqrt(Es/N0)) within 5 standard deviations, the FER falls with
    Es/N0, the counts are consistent, and a seed repeats the curve."""
    code = seeded_code("r1/2-z27")
    c = Constellation.qam(4)
    wf = waterfall_db("r1/2-z27", 4)
    grid = [wf - 1.0, wf + 0.5]
    curve = error_rate_curve(code, constellation=c, esn0_db=grid, min_block_errors=40,
                             max_blocks=3000, rng=5)  # fmt: skip
    assert np.all(curve.num_blocks >= 1) and np.all(curve.num_blocks <= 3000)
    for i, e in enumerate(grid):
        p = 0.5 * erfc(math.sqrt(10 ** (e / 10)) / math.sqrt(2))
        bits = curve.num_blocks[i] * curve.n
        assert abs(curve.pre_fec_ber[i] - p) < 5 * math.sqrt(p / bits)
    assert curve.fer[1] < curve.fer[0]
    assert np.all(curve.bit_errors >= curve.block_errors)
    assert np.all(curve.undetected_errors <= curve.block_errors)
    assert np.all(curve.mean_iterations >= 1)
    again = error_rate_curve(code, constellation=c, esn0_db=grid, min_block_errors=40,
                             max_blocks=3000, rng=5)  # fmt: skip
    np.testing.assert_array_equal(again.bit_errors, curve.bit_errors)
    np.testing.assert_array_equal(again.total_iterations, curve.total_iterations)


def test_error_rate_curve_of_pas():
    c = Constellation.qam(256).shaped(nu=0.05)
    pas = PAS(seeded_code("r5/6-z360"), c)
    curve = error_rate_curve(pas, constellation=c, esn0_db=[16.0], max_blocks=64, rng=1)
    assert curve.num_blocks[0] == 64 and curve.k == pas.k
    assert curve.block_errors[0] == 0
    with pytest.raises(ValueError, match="another constellation"):
        error_rate_curve(pas, constellation=Constellation.qam(256), esn0_db=[16.0])


@pytest.mark.parametrize(("name", "order"), [("r4/5-z360", 64), ("r1/2-z27", 16)])
def test_threshold_lies_below_the_measured_waterfall(name, order):
    """The long-code sum-product prediction is optimistic by well under 3 dB
    against min-sum on these codeword lengths."""
    pred = threshold_esn0_db(seeded_code(name), Constellation.qam(order))
    assert pred < waterfall_db(name, order) < pred + 3.0
