"""Standard LDPC codes on the host: DVB-S2 (ETSI EN 302 307-1) and the
25G/50G-PON code (ITU-T G.9804.2), checked against the standards'
dimensions and degree distributions, their codewords, and their EXIT
thresholds (the reference's tests/coding/test_ldpc_tables.py and
test_ldpc_exit.py)."""

import numpy as np
import pytest

from commkit_amd import LDPCCode
from commkit_amd.coding.exit import ldpc_threshold

from .codes import encode_host, syndrome

# k of each rate label: normal frames n * rate; short frames from Table 5b
# of the standard (five labels carry fewer bits than 16200 * rate).
K_NORMAL = {"1/4": 16200, "1/3": 21600, "2/5": 25920, "1/2": 32400, "3/5": 38880,
            "2/3": 43200, "3/4": 48600, "4/5": 51840, "5/6": 54000, "8/9": 57600,
            "9/10": 58320}  # fmt: skip
K_SHORT = {"1/4": 3240, "1/3": 5400, "2/5": 6480, "1/2": 7200, "3/5": 9720, "2/3": 10800,
           "3/4": 11880, "4/5": 12600, "5/6": 13320, "8/9": 14400}  # fmt: skip


@pytest.mark.parametrize(("frame", "rate"), [("normal", r) for r in K_NORMAL]
                         + [("short", r) for r in K_SHORT])  # fmt: skip
def test_dvbs2_dimensions_and_accumulator(frame, rate):
    code = LDPCCode.dvbs2(rate, frame=frame)
    n = 64800 if frame == "normal" else 16200
    k = (K_NORMAL if frame == "normal" else K_SHORT)[rate]
    assert (code.n, code.k, code.num_checks) == (n, k, n - k)
    assert code._encoder.kind == "accumulate"
    degree = np.bincount(code.variables, minlength=n)
    np.testing.assert_array_equal(degree[k:-1], 2)  # parity: degree 2, the last 1
    assert degree[-1] == 1
    info = degree[:k].reshape(-1, 360)  # whole groups of 360, two degrees
    assert np.all(info == info[:, :1]) and np.unique(info).size <= 2


@pytest.mark.parametrize(("rate", "high", "count"), [("1/2", 8, 12960), ("3/4", 12, 5400)])
def test_dvbs2_degree_distribution(rate, high, count):
    """Table 7a of the standard: the bits of the high degree; the rest 3."""
    code = LDPCCode.dvbs2(rate)
    degree = np.bincount(code.variables, minlength=64800)[: code.k]
    assert np.count_nonzero(degree == high) == count
    assert np.count_nonzero(degree == 3) == code.k - count


def test_dvbs2_codewords():
    code = LDPCCode.dvbs2("2/3", frame="short")
    u = np.random.default_rng(0).integers(0, 2, (3, code.k))
    assert not syndrome(code, encode_host(code, u)).any()


def test_pon_structure():
    """ITU-T G.9804.2 Annex B.1: 12 x 69 circulants of 256, 384 punctured."""
    code = LDPCCode.pon()
    assert (code.num_variables, code.num_checks, code.punctured) == (17664, 3072, 384)
    assert (code.n, code.k) == (17280, 14592)
    degree = np.bincount(code.variables, minlength=17664).reshape(69, 256)
    assert np.all(degree == degree[:, :1])
    np.testing.assert_array_equal(degree[:, 0], [6] * 17 + [3] * 50 + [12, 11])
    rows = np.bincount(code.checks, minlength=3072).reshape(12, 256)
    assert np.all(rows == rows[:, :1])
    assert code._qc_view is not None  # decoded by the quasi-cyclic kernels


def test_dvbs2_thresholds():
    """The rate-1/2 short code opens its tunnel at an information per bit of
    0.49-0.505 (where QPSK frames start to decode), the normal code above
    but near its rate."""
    assert 0.49 < ldpc_threshold(LDPCCode.dvbs2("1/2", frame="short")) < 0.505
    code = LDPCCode.dvbs2("1/2")
    assert code.k / code.n < ldpc_threshold(code) < code.k / code.n + 0.05


def test_validation():
    with pytest.raises(ValueError, match="frame"):
        LDPCCode.dvbs2("1/2", frame="medium")
    with pytest.raises(ValueError, match="rate must be one of"):
        LDPCCode.dvbs2("9/10", frame="short")
