"""EXIT analysis on the host: J against adaptive quadrature, the (3, 6)
ensemble's threshold against its known density-evolution value, and the
curve identities of the reference's tests/coding/test_ldpc_exit.py."""

import math

import numpy as np
import pytest
from scipy import integrate

from commkit_amd import LDPCCode
from commkit_amd.coding.exit import (
    ExitCurve,
    _j,
    _j_inv,
    exit_chart,
    ldpc_exit_curves,
    ldpc_threshold,
)
from commkit_amd.coding.performance import ErrorRateCurve
from commkit_amd.workloads import CODES, seeded_code


def _qc(z, shifts, punctured=0):
    """Quasi-cyclic code from its edges: one z x z circulant per entry."""
    rows, cols = [], []
    for i, row in enumerate(shifts):
        for j, s in enumerate(row):
            rows += [i * z + t for t in range(z)]
            cols += [j * z + (t + s) % z for t in range(z)]
    return LDPCCode(np.array(rows), np.array(cols), shifts.shape[0] * z, shifts.shape[1] * z,
                    punctured=punctured)  # fmt: skip


REGULAR_3_6 = _qc(100, np.random.default_rng(0).integers(0, 100, (3, 6)))


def _j_quad(sigma):
    """1 - E[log2(1 + e^-L)], L ~ N(sigma^2/2, sigma^2), adaptive quadrature."""
    mu, s2 = sigma * sigma / 2, sigma * sigma

    def f(v):
        return (np.exp(-((v - mu) ** 2) / (2 * s2)) / math.sqrt(2 * math.pi * s2)
                * np.logaddexp(0.0, -v) / math.log(2.0))  # fmt: skip

    return 1.0 - integrate.quad(f, mu - 40 * sigma, mu + 40 * sigma, limit=500, epsabs=1e-14,
                                epsrel=1e-12)[0]  # fmt: skip


def _bpsk_mi(ebn0_db, rate):
    """Information of BPSK-AWGN channel LLRs: J(sqrt(8 R Eb/N0))."""
    return _j_quad(math.sqrt(8 * rate * 10 ** (ebn0_db / 10)))


def test_j_against_quadrature():
    sig = np.linspace(0.01, 14.0, 60)
    want = np.array([_j_quad(s) for s in sig])
    np.testing.assert_allclose(_j(sig), want, rtol=0, atol=1e-7)
    # J^-1 is ill-conditioned where J is within 1e-4 of 1 (sigma above ~7)
    low = sig < 7.0
    np.testing.assert_allclose(_j_inv(want[low]), sig[low], rtol=2e-5, atol=0)
    assert _j(0.0) == 0.0 and _j(40.0) == pytest.approx(1.0, abs=1e-13)


def test_regular_3_6_threshold():
    """Density evolution puts the (3, 6) ensemble at 1.11 dB Eb/N0; the
    Gaussian approximation lands within 0.1 dB."""
    thr = ldpc_threshold(REGULAR_3_6)
    assert _bpsk_mi(1.01, 0.5) < thr < _bpsk_mi(1.21, 0.5)


def test_chart_converges_above_the_threshold_only():
    thr = ldpc_threshold(REGULAR_3_6)
    above = exit_chart(*ldpc_exit_curves(REGULAR_3_6, channel_mi=thr + 0.005))
    below = exit_chart(*ldpc_exit_curves(REGULAR_3_6, channel_mi=thr - 0.005))
    assert above.converges and not below.converges
    assert above.trajectory[-1, 0] >= 0.99 and below.trajectory[-1, 0] < 0.99
    assert below.iterations > above.iterations or below.trajectory[-1, 0] < 0.99


def test_without_a_priori_the_variable_nodes_pass_the_channel():
    for mi in (0.2, 0.5, 0.9):
        variable, _ = ldpc_exit_curves(REGULAR_3_6, channel_mi=mi)
        assert variable.extrinsic[0] == pytest.approx(mi, abs=2e-3)


def test_punctured_variables_get_no_channel():
    code = _qc(50, np.random.default_rng(1).integers(0, 50, (3, 6)), punctured=50)
    variable, _ = ldpc_exit_curves(code, channel_mi=0.5)
    seen = np.isin(code.variables, np.arange(code.n))
    assert variable.extrinsic[0] == pytest.approx(0.5 * seen.mean(), abs=2e-3)


def test_degree_two_checks_are_the_identity():
    code = _qc(40, np.array([[0, 3]]))
    _, check = ldpc_exit_curves(code, channel_mi=0.7)
    np.testing.assert_allclose(check.extrinsic, check.a_priori, atol=2e-3)


def test_endpoints_labels_and_trajectory():
    variable, check = ldpc_exit_curves(REGULAR_3_6, channel_mi=0.5, num_points=11)
    assert variable.a_priori.shape == (11,)
    assert check.extrinsic[0] == pytest.approx(0.0, abs=1e-6)
    assert check.extrinsic[-1] == pytest.approx(1.0, abs=1e-6)
    assert variable.extrinsic[-1] == pytest.approx(1.0, abs=1e-6)
    assert np.all(np.diff(variable.extrinsic) >= 0) and np.all(np.diff(check.extrinsic) >= 0)
    assert variable.label == "variable nodes, I_ch = 0.500" and check.label == "check nodes"
    variable, check = ldpc_exit_curves(REGULAR_3_6, channel_mi=0.6)
    steps = exit_chart(variable, check).trajectory[1:]
    up, across = steps[0::2], steps[1::2]
    np.testing.assert_allclose(up[:, 1], np.interp(up[:, 0], variable.a_priori,
                                                   variable.extrinsic))  # fmt: skip
    np.testing.assert_allclose(across[:, 0], np.interp(across[:, 1], check.a_priori,
                                                       check.extrinsic))  # fmt: skip


@pytest.mark.parametrize("name", sorted(CODES))
def test_seeded_codes_open_above_their_rate(name):
    """Every benchmark code opens its tunnel above its rate: within 0.07 of
    it for the weighted bases, within 0.15 for the random-density rate-1/2
    base (0.633 measured)."""
    code = seeded_code(name)
    thr = ldpc_threshold(code)
    gap = 0.15 if name == "r1/2-z27" else 0.07
    assert code.k / code.n < thr < code.k / code.n + gap


def test_validation():
    with pytest.raises(TypeError, match="LDPCCode"):
        ldpc_threshold("not a code")
    with pytest.raises(TypeError, match="LDPCCode"):
        ldpc_exit_curves(None, channel_mi=0.5)
    for mi in (0.0, 1.0, -0.1):
        with pytest.raises(ValueError, match="channel_mi"):
            ldpc_exit_curves(REGULAR_3_6, channel_mi=mi)
    with pytest.raises(ValueError, match="num_points"):
        ldpc_exit_curves(REGULAR_3_6, channel_mi=0.5, num_points=1)
    with pytest.raises(ValueError, match="increase"):
        ExitCurve([0.0, 0.0], [0.1, 0.2])
    with pytest.raises(ValueError, match="span"):
        exit_chart(ExitCurve([0.0, 0.5], [0.1, 0.2]), ExitCurve([0.0, 1.0], [0.0, 1.0]))


def test_error_rate_curve_record():
    """The counts record of the Monte-Carlo curves: rates and validation."""
    curve = ErrorRateCurve([1.0, 2.0], num_blocks=[10, 20], bit_errors=[30, 1],
                           block_errors=[5, 1], pre_fec_bit_errors=[100, 50],
                           detected_failures=[5, 1], undetected_errors=[0, 0],
                           total_iterations=[200, 100], k=8, n=16, bits_per_symbol=2)  # fmt: skip
    np.testing.assert_allclose(curve.fer, [0.5, 0.05])
    np.testing.assert_allclose(curve.ber, [30 / 80, 1 / 160])
    np.testing.assert_allclose(curve.pre_fec_ber, [100 / 160, 50 / 320])
    np.testing.assert_allclose(curve.mean_iterations, [20, 5])
    np.testing.assert_allclose(curve.ebn0_db, [1.0, 2.0])  # rate 1/2 x 2 bits: Eb = Es
    with pytest.raises(ValueError, match="num_blocks"):
        ErrorRateCurve([1.0], [0], [0], [0], [0], [0], [0], [0], k=1, n=2, bits_per_symbol=1)
    with pytest.raises(ValueError, match="shape"):
        ErrorRateCurve([1.0], [1, 2], [0], [0], [0], [0], [0], [0], k=1, n=2, bits_per_symbol=1)
