"""Scoring arguments checked on the host, before any device work."""

import numpy as np
import pytest

from commkit_amd import Constellation, DecodeResult, Reference, Signal
from commkit_amd.metrics import ber, fer, gmi, ngmi
from commkit_amd.workloads import seeded_code


def _result(bits):
    return DecodeResult(bits=bits, llrs=None, valid=np.ones(1, bool),
                        iterations=np.ones(1, np.int32), checked=True)  # fmt: skip


def test_argument_errors():
    bits = np.zeros(16, np.int8)
    with pytest.raises(ValueError, match="reference"):
        ber(bits)
    with pytest.raises(ValueError, match="num_skip_symbols"):
        ber(_result(bits), bits, num_skip_symbols=1)
    with pytest.raises(ValueError, match="constellation"):
        ber(bits, bits, num_skip_symbols=1)
    with pytest.raises(ValueError, match="info_bits"):
        fer(bits, bits)
    with pytest.raises(ValueError, match="disagrees"):
        fer(_result(bits), Reference(bits=bits, code=seeded_code("r1/2-z27")), info_bits=8)
    with pytest.raises(ValueError, match="DecodeResult needs"):
        fer(_result(bits), None)
    with pytest.raises(ValueError, match="no bits"):
        fer(_result(bits), Reference(symbols=np.zeros(4)))
    with pytest.raises(ValueError, match="constellation"):
        gmi(np.zeros(8, np.float32), bits[:8])
    with pytest.raises(ValueError, match="noise_var"):
        ngmi(np.zeros(8, np.float32), bits[:8], constellation=Constellation.qam(4), noise_var=1.0)
    with pytest.raises(ValueError, match="noise_var or esn0_db"):
        gmi(Signal(np.zeros(4, np.complex64), constellation=Constellation.qam(4)))
    with pytest.raises(TypeError):
        fer(bits, bits, info_bits=2.0)
