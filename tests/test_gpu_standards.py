"""Standard codes on the GPU: DVB-S2 through the generic (CSR) kernels, APP
in LDS (short frames) and in HBM (normal frames), bit-exact min-sum against
the oracle; the waterfall of the rate-1/2 normal code and the PON code's
correction of a 2e-2 pre-FEC BER, through generate, apply_awgn, compute_llr,
decode and the metrics (the reference's tests/coding/test_ldpc.py:262-311)."""

import math

import numpy as np
import pytest
from scipy.special import erfc

from commkit_amd import Constellation, LDPCCode, MinSum, SumProduct, compute_llr, decode
from commkit_amd.core import generate
from commkit_amd.impairments import apply_awgn
from commkit_amd.metrics import ber, fer

from .test_gpu_ldpc import assert_identical, gpu_decode, oracle_decode

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize(("frame", "rows"), [("short", 8), ("normal", 2)])
@pytest.mark.parametrize("sum_product", [False, True])
def test_dvbs2_matches_the_oracle(frame, rows, sum_product):
    """Min-sum bit for bit (APP, bits, iterations, valid); sum-product to
    the reference's tolerance, frames near the waterfall so that the
    iteration counts spread."""
    code = LDPCCode.dvbs2("1/2", frame=frame)
    r = np.random.default_rng(3)
    llrs = (1.1 + 1.5 * r.standard_normal((rows, code.n))).astype(np.float32)
    method = (SumProduct(max_iterations=25) if sum_product
              else MinSum(scale=0.875, max_iterations=25))  # fmt: skip
    got, want = gpu_decode(code, llrs, method), oracle_decode(code, llrs, method)
    if not sum_product:
        assert_identical(got, want, f"dvbs2 {frame}")
    else:
        same = (got[3] == want[3]) & (got[0] == want[0]).all(axis=1)
        assert same.mean() >= 0.75
        np.testing.assert_allclose(got[1][same], want[1][same], rtol=1e-3, atol=1e-3)
    assert len(set(want[3].tolist())) > 1 or not want[2].all()


def _qpsk_ber(esn0):
    return 0.5 * erfc(math.sqrt(esn0) / math.sqrt(2))


@pytest.mark.parametrize("method", [MinSum(), SumProduct()], ids=["min-sum", "sum-product"])
def test_dvbs2_waterfall(method):
    """QPSK, rate 1/2, normal frames (the standard quotes Es/N0 = 1.0 dB at a
    packet error rate of 1e-7 with BCH): 12 frames all decode 0.3 dB above
    it, most fail 0.6 dB below it; the channel's pre-FEC BER is Gray QPSK's
    Q(sqrt(Es/N0)) within 4 standard deviations first."""
    code = LDPCCode.dvbs2("1/2")
    c = Constellation.qam(4)
    frames = 12
    sig = generate(c, num_symbols=32400 * frames, symbol_rate=1e9, code=code, rng=6,
                   device="gpu")  # fmt: skip
    bits = 64800 * frames
    for es, check in ((1.3, lambda f: f == 0.0), (0.4, lambda f: f > 0.5)):
        rx = apply_awgn(sig, esn0_db=es, rng=7)
        p = _qpsk_ber(10 ** (es / 10))
        assert abs(ber(rx) - p) < 4 * math.sqrt(p * (1 - p) / bits)
        res = decode(compute_llr(rx, noise_var=10 ** (-es / 10)), code=code, method=method)
        f = fer(res.bits, sig.reference.info_bits, info_bits=code.k)
        assert check(f), (es, f)


def test_pon_corrects_its_threshold():
    """LDPC(17280, 14592), quasi-cyclic with punctured parity: soft decoding
    clears a pre-FEC BER near 2e-2."""
    code = LDPCCode.pon()
    sig = generate(Constellation.qam(4), num_symbols=8640 * 8, symbol_rate=1e9, code=code,
                   rng=8, device="gpu")  # fmt: skip
    rx = apply_awgn(sig, esn0_db=6.0, rng=9)
    assert 0.02 < ber(rx) < 0.025
    res = decode(compute_llr(rx, noise_var=10**-0.6), code=code)
    assert ber(res.bits, sig.reference.info_bits) == 0.0
    assert bool(res.valid.all())
