"""The standard LDPC codes against their standards (host).

DVB-S2 (ETSI EN 302 307-1, clause 5.3.2 and Annexes B/C): every normal and
short code is checked against one table, per code its information length
(Tables 5a/5b), its parity-group parameter q = m / 360 and its
information-bit degree profile (Tables 7a/7b): ``high`` bits of the high
degree, first in the frame, then degree 3.  The parity part must be the
accumulator (H[:, k:] lower bidiagonal), and the information bits must keep
the code's group structure: bit 360 g + j checks the rows of bit 360 g
shifted by j q (mod m).  The 25G/50G-PON code (ITU-T G.9804.2 Annex B.1)
is a 12 x 69 base matrix of circulants of 256 with 384 parity bits
punctured.  Codewords of the library's encoder satisfy H c = 0; the EXIT
thresholds of the rate-1/2 codes sit where the reference's test puts them
(tests/coding/test_ldpc_exit.py of the reference).
"""

import numpy as np
import pytest

from commkit_amd import LDPCCode
from commkit_amd.coding.exit import ldpc_threshold

from .codes import encode_host, syndrome

# (frame, rate) -> (k, q, high degree, bits of the high degree)
DVBS2 = {
    ("normal", "1/4"): (16200, 135, 12, 5400),
    ("normal", "1/3"): (21600, 120, 12, 7200),
    ("normal", "2/5"): (25920, 108, 12, 8640),
    ("normal", "1/2"): (32400, 90, 8, 12960),
    ("normal", "3/5"): (38880, 72, 12, 12960),
    ("normal", "2/3"): (43200, 60, 13, 4320),
    ("normal", "3/4"): (48600, 45, 12, 5400),
    ("normal", "4/5"): (51840, 36, 11, 6480),
    ("normal", "5/6"): (54000, 30, 13, 5400),
    ("normal", "8/9"): (57600, 20, 4, 7200),
    ("normal", "9/10"): (58320, 18, 4, 6480),
    ("short", "1/4"): (3240, 36, 12, 1440),
    ("short", "1/3"): (5400, 30, 12, 1800),
    ("short", "2/5"): (6480, 27, 12, 2160),
    ("short", "1/2"): (7200, 25, 8, 1800),
    ("short", "3/5"): (9720, 18, 12, 3240),
    ("short", "2/3"): (10800, 15, 13, 1080),
    ("short", "3/4"): (11880, 12, 12, 360),
    ("short", "4/5"): (12600, 10, 3, 12600),
    ("short", "5/6"): (13320, 8, 13, 360),
    ("short", "8/9"): (14400, 5, 4, 1800),
}  # fmt: skip


@pytest.mark.parametrize("frame,rate", list(DVBS2), ids=[f"{f}-{r}" for f, r in DVBS2])
def test_dvbs2_against_the_standard(frame, rate):
    k, q, high, count = DVBS2[(frame, rate)]
    code = LDPCCode.dvbs2(rate, frame=frame)
    n = {"normal": 64800, "short": 16200}[frame]
    m = 360 * q
    assert (code.n, code.k, code.num_checks, code.punctured) == (n, k, m, 0)
    h = code.parity_check().tocsc()
    # parity part: ones on the diagonal and just below it, nothing else
    hp = h[:, k:].tocoo()
    assert hp.nnz == 2 * m - 1
    assert np.all((hp.row == hp.col) | (hp.row == hp.col + 1))
    assert code._encoder.kind == "accumulate"
    # information bits: the standard's degree profile, high degrees first
    deg = np.diff(h.indptr)[:k]
    assert np.all(deg[:count] == high) and np.all(deg[count:] == 3)
    # group structure: bit 360 g + j checks the rows of bit 360 g, + j q mod m
    for g in (0, k // 360 - 1):
        first = np.sort(h.indices[h.indptr[360 * g] : h.indptr[360 * g + 1]])
        for j in (1, 7, 359):
            b = 360 * g + j
            rows = np.sort(h.indices[h.indptr[b] : h.indptr[b + 1]])
            np.testing.assert_array_equal(rows, np.sort((first + j * q) % m))


@pytest.mark.parametrize("frame,rate", [("short", "2/3"), ("short", "1/4"), ("normal", "5/6")])
def test_dvbs2_encoder_codewords(frame, rate):
    code = LDPCCode.dvbs2(rate, frame=frame)
    u = np.random.default_rng(len(rate)).integers(0, 2, (2, code.k))
    words = encode_host(code, u)
    assert not syndrome(code, words).any()
    np.testing.assert_array_equal(words[:, : code.k], u)  # systematic


def test_pon_against_the_standard():
    """ITU-T G.9804.2 Annex B.1: circulants of 256 on a 12 x 69 base matrix;
    column weights 6 (17 columns), 3 (50), 12 and 11; 384 parity bits
    punctured, so n = 17280 and k = 14592."""
    code = LDPCCode.pon()
    base, z = code.__dict__["_qc"]
    assert z == 256 and base.shape == (12, 69)
    present = base >= 0
    assert base[present].max() < 256
    np.testing.assert_array_equal(present.sum(axis=0), [6] * 17 + [3] * 50 + [12, 11])
    assert (code.num_variables, code.num_checks, code.punctured) == (69 * 256, 12 * 256, 384)
    assert (code.n, code.k) == (17280, 14592)
    # every nonzero block of H is the identity rolled by its shift
    h = code.parity_check().tocsr()
    for i, j in ((0, 0), (5, 20), (11, 68)):
        if present[i, j]:
            block = h[256 * i : 256 * (i + 1), 256 * j : 256 * (j + 1)].toarray()
            np.testing.assert_array_equal(block, np.roll(np.eye(256, dtype=block.dtype),
                                                         base[i, j], axis=1))  # fmt: skip
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
