"""Signal archives on the host: save_npz / load_npz round trips, the codes
rebuilt from their fields (quasi-cyclic structure recovered), and the
refusals (pickled metadata, frames, pulses).  Mirrors the archive tests of
the reference's tests/test_io.py for the fields the receive tail reads."""

import json

import numpy as np
import pytest

from commkit_amd import Constellation, LDPCCode, Reference, Signal
from commkit_amd.coding import CCDM, PAS
from commkit_amd.io import _quasi_cyclic, load_npz, save_npz
from commkit_amd.workloads import seeded_base, seeded_code

PS256 = Constellation.qam(256).shaped(nu=0.05)


def _same_code(a, b):
    assert type(a) is type(b)
    if isinstance(a, LDPCCode):
        np.testing.assert_array_equal(a.checks, b.checks)
        np.testing.assert_array_equal(a.variables, b.variables)
        assert (a.num_checks, a.num_variables, a.punctured) == (b.num_checks, b.num_variables,
                                                                  b.punctured)  # fmt: skip
    elif isinstance(a, CCDM):
        assert a == b
    else:
        _same_code(a.code, b.code)
        assert a.constellation == b.constellation


@pytest.mark.parametrize("kind", ["ldpc", "ccdm", "pas", "uncoded"])
def test_round_trip(tmp_path, kind):
    rng = np.random.default_rng(0)
    c = Constellation.qam(16)
    if kind == "ldpc":
        code = seeded_code("r2/3-z96")
    elif kind == "ccdm":
        code = CCDM(Constellation(np.array([1.0, 3.0, 5.0, 7.0]), pmf=[0.4, 0.3, 0.2, 0.1]),
                    num_symbols=40)  # fmt: skip
    elif kind == "pas":
        code, c = PAS(seeded_code("r5/6-z360"), PS256), PS256
    else:
        code = None
    bits = rng.integers(0, 2, size=(2, 64), dtype=np.int8)
    info = rng.integers(0, 2, size=(2, 40), dtype=np.int8) if code is not None else bits
    sym = (rng.standard_normal((2, 16)) + 1j * rng.standard_normal((2, 16))).astype(np.complex64)
    sig = Signal(sym + 0.1, constellation=c,
                 reference=Reference(symbols=sym, bits=bits, info_bits=info, code=code))  # fmt: skip
    save_npz(sig, tmp_path / "cap")  # ".npz" appended
    back = load_npz(tmp_path / "cap.npz")
    np.testing.assert_array_equal(back.samples, sig.samples)
    assert back.samples.dtype == np.complex64 and back.sps == 1 and back.frame is None
    assert back.constellation == c and back.constellation.family == c.family
    ref = back.reference
    np.testing.assert_array_equal(ref.symbols, sym)
    np.testing.assert_array_equal(ref.bits, bits)
    np.testing.assert_array_equal(ref.info_bits, info)
    if code is None:
        assert ref.code is None
    else:
        _same_code(ref.code, code)


def test_qc_code_keeps_its_base(tmp_path):
    code = seeded_code("r5/6-z360")
    sig = Signal(np.zeros(4, np.complex64), constellation=Constellation.qam(64),
                 reference=Reference(bits=np.zeros(code.n, np.int8), code=code))  # fmt: skip
    save_npz(sig, tmp_path / "qc.npz", compressed=False)
    loaded = load_npz(tmp_path / "qc.npz").reference.code
    base, z = loaded.__dict__["_qc"]
    assert z == 360
    np.testing.assert_array_equal(base, np.where(seeded_base("r5/6-z360") >= 0,
                                                 seeded_base("r5/6-z360") % 360, -1))  # fmt: skip


def test_quasi_cyclic_detection():
    """A QC code given by its edges gets its base back; the largest lifting
    wins (a Z = 2 code is also one of Z = 1 blocks, which is not a lifting);
    a code that is not made of circulants comes back as it was."""
    code = seeded_code("r1/2-z27")
    plain = LDPCCode(code.checks, code.variables, code.num_checks, code.num_variables,
                     code.punctured)  # fmt: skip
    assert "_qc" not in plain.__dict__
    qc = _quasi_cyclic(plain)
    assert qc.__dict__["_qc"][1] == 27
    np.testing.assert_array_equal(qc.__dict__["_qc"][0], code.__dict__["_qc"][0])
    h = np.array([[1, 1, 0, 1, 0, 0], [0, 1, 1, 0, 1, 0], [1, 0, 1, 0, 0, 1]])
    irregular = LDPCCode.from_parity_check(h)
    assert _quasi_cyclic(irregular) is irregular


def test_gpu_tensors_are_saved_from_the_host_copy(tmp_path):
    import torch

    t = torch.arange(8, dtype=torch.float32)  # a CPU tensor stands in for a device one
    save_npz(Signal(t, constellation=Constellation.pam(4)), tmp_path / "t.npz")
    np.testing.assert_array_equal(load_npz(tmp_path / "t.npz").samples, t.numpy())


def test_refusals(tmp_path):
    with pytest.raises(TypeError, match="Signal"):
        save_npz(np.zeros(3), tmp_path / "x.npz")
    with pytest.raises(ValueError, match="framed"):
        save_npz(Signal(np.zeros(3), frame=object()), tmp_path / "x.npz")
    with pytest.raises(ValueError, match="device"):
        load_npz(tmp_path / "x.npz", device="tpu")
    # Pickled metadata is never unpickled.
    np.savez(tmp_path / "old.npz", samples=np.zeros(3),
             __metadata__=np.array([{"a": 1}], dtype=object))  # fmt: skip
    with pytest.raises(ValueError):
        load_npz(tmp_path / "old.npz")
    meta = {"sampling_rate": 2.0, "symbol_rate": 1.0, "center_frequency": 0.0,
            "domain": "time", "pulse": {"type": "RRC", "rolloff": 0.1}}  # fmt: skip
    np.savez(tmp_path / "pulse.npz", samples=np.zeros(3), __metadata__=np.array(json.dumps(meta)))
    with pytest.raises(ValueError, match="pulse"):
        load_npz(tmp_path / "pulse.npz")
    meta["pulse"] = None
    np.savez(tmp_path / "frame.npz", samples=np.zeros(3), __metadata__=np.array(json.dumps(meta)),
             __frame_metadata__=np.array("{}"))  # fmt: skip
    with pytest.raises(ValueError, match="framed"):
        load_npz(tmp_path / "frame.npz")
    meta["reference_code"] = {"type": "Interleaver", "fields": {}}
    np.savez(tmp_path / "code.npz", samples=np.zeros(3), reference_bits=np.zeros(3, np.int8),
             __metadata__=np.array(json.dumps(meta)))  # fmt: skip
    with pytest.raises(ValueError, match="Unknown code type"):
        load_npz(tmp_path / "code.npz")


def test_oversampled_record_keeps_its_rate(tmp_path):
    meta = {"sampling_rate": 4e9, "symbol_rate": 2e9, "center_frequency": 0.0,
            "domain": "time", "pulse": None, "constellation_family": None,
            "reference_code": None}  # fmt: skip
    np.savez(tmp_path / "os.npz", samples=np.zeros(8, np.complex64),
             __metadata__=np.array(json.dumps(meta)))  # fmt: skip
    sig = load_npz(tmp_path / "os.npz")
    assert sig.sps == 2 and sig.constellation is None and sig.reference is None
