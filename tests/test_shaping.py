"""Distribution matching on the host: the CCDM tables and the oracle's range
coder, pinned by brute force and a hand-worked block (no GPU).

Mirrors tests/coding/test_shaping.py of the reference (composition,
exhaustive round trip, validity exactly on the image, two points of four
sequences, rate, validation, PAS layout)."""

import itertools

import numpy as np
import pytest

from commkit_amd import Constellation
from commkit_amd.coding import CCDM, PAS
from commkit_amd.coding.shaping import _composition
from commkit_amd.workloads import seeded_code
from oracle import cpu

AMPLITUDES = Constellation(np.array([1.0, 3.0, 5.0, 7.0]), pmf=[0.4, 0.3, 0.2, 0.1])


def _kl(c: np.ndarray, p: np.ndarray) -> float:
    q = c / c.sum()
    nz = q > 0
    return float(np.sum(q[nz] * np.log(q[nz] / p[nz])))


@pytest.mark.parametrize("n", [5, 9, 16])
def test_composition_minimizes_the_divergence(n):
    """Against every composition of n points of four kinds."""
    p = np.asarray(AMPLITUDES.pmf)
    best = min((c for c in itertools.product(range(n + 1), repeat=4) if sum(c) == n),
               key=lambda c: _kl(np.array(c, float), p))  # fmt: skip
    got = _composition(p, n)
    assert got.sum() == n
    assert _kl(got.astype(float), p) == pytest.approx(_kl(np.array(best, float), p))


def test_hand_worked_block():
    """Composition (2, 1), n = 3, k = 1, worked by hand from the coder's
    definition.  Position 0 splits 2^32 into floor(2^33/3) = 2863311530 and
    floor(2^32/3) = 1431655765, the remainder 1 to point 0 (count 2):
    [2863311531, 1431655765].  Payload 0 (fraction 0) stays in point 0; the
    range 2863311531 splits evenly into 1431655765 + 1431655765 with the
    remainder to point 0 (the first of equal counts): [1431655766,
    1431655765], so 0 again, and the last point is 1: (0, 0, 1).  Payload 1
    (fraction 2^31) is point 0, then 2^31 >= 1431655766 makes point 1, then
    point 0: (0, 1, 0).  (1, 0, 0) is a sequence of the composition that no
    payload reaches."""
    c = CCDM(Constellation(np.array([1.0, 3.0]), pmf=[2 / 3, 1 / 3]), num_symbols=3)
    np.testing.assert_array_equal(c.composition, [2, 1])
    assert c.k == 1
    sym = cpu.ccdm_match(np.array([[0], [1]], np.int8), c.composition)
    np.testing.assert_array_equal(sym, [[0, 0, 1], [0, 1, 0]])
    bits, valid = cpu.ccdm_dematch(np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]]), c.composition,
                                   c.k, c._buffer_bits)  # fmt: skip
    np.testing.assert_array_equal(bits[:2], [[0], [1]])
    np.testing.assert_array_equal(valid, [True, True, False])


PMFS = [[0.4, 0.3, 0.2, 0.1], [0.25] * 4, [0.7, 0.2, 0.05, 0.05]]


@pytest.mark.parametrize("n", [3, 4, 6, 8, 11])
@pytest.mark.parametrize("pmf", PMFS)
def test_oracle_exhaustive_round_trip(n, pmf):
    """Every k-bit payload: the exact composition, distinct sequences in the
    payloads' order (an arithmetic coder preserves order), and the inverse."""
    cc = CCDM(Constellation(AMPLITUDES.points, pmf=pmf), num_symbols=n)
    u = np.array(list(itertools.product([0, 1], repeat=cc.k)), dtype=np.int8)
    sym = cpu.ccdm_match(u, cc.composition)
    for row in sym:
        np.testing.assert_array_equal(np.bincount(row, minlength=4), cc.composition)
    keys = [tuple(r) for r in sym]
    assert len(set(keys)) == len(keys)
    assert keys == sorted(keys)
    back, valid = cpu.ccdm_dematch(sym, cc.composition, cc.k, cc._buffer_bits)
    np.testing.assert_array_equal(back, u)
    assert valid.all()


@pytest.mark.parametrize("n", [3, 4, 6, 8])
@pytest.mark.parametrize("pmf", PMFS)
def test_oracle_valid_exactly_on_the_image(n, pmf):
    cc = CCDM(Constellation(AMPLITUDES.points, pmf=pmf), num_symbols=n)
    u = np.array(list(itertools.product([0, 1], repeat=cc.k)), dtype=np.int8)
    image = {tuple(r) for r in cpu.ccdm_match(u, cc.composition)}
    every = np.array(list(itertools.product(range(4), repeat=n)), dtype=np.int32)
    every = every[(np.stack([(every == i).sum(axis=1) for i in range(4)], axis=1)
                   == cc.composition).all(axis=1)]  # fmt: skip
    assert len(every) > len(image)
    _, valid = cpu.ccdm_dematch(every, cc.composition, cc.k, cc._buffer_bits)
    np.testing.assert_array_equal(valid, [tuple(r) in image for r in every])


def test_two_points_of_four_sequences():
    """Two of each of two points: six sequences, four payloads (k = 2)."""
    cc = CCDM(Constellation.pam(2), num_symbols=4)
    assert cc.k == 2
    every = np.array([r for r in itertools.product(range(2), repeat=4) if sum(r) == 2])
    _, valid = cpu.ccdm_dematch(every, cc.composition, cc.k, cc._buffer_bits)
    assert valid.sum() == 4


def test_oracle_round_trip_at_operating_size():
    """The amplitude matcher of shaped 256-QAM on a 21,600-bit codeword."""
    pas = PAS(seeded_code("r5/6-z360"), Constellation.qam(256).shaped(nu=0.05))
    cc = pas.matcher
    u = np.random.default_rng(1).integers(0, 2, size=(4, cc.k), dtype=np.int8)
    sym = cpu.ccdm_match(u, cc.composition)
    for row in sym:
        np.testing.assert_array_equal(np.bincount(row, minlength=8), cc.composition)
    back, valid = cpu.ccdm_dematch(sym, cc.composition, cc.k, cc._buffer_bits)
    np.testing.assert_array_equal(back, u)
    assert valid.all()
    # A point changing kind breaks the composition.  Two points trading places
    # keep it: the image holds over half of those sequences (log2 of their
    # number exceeds k by under one bit here), so the swap is either flagged or is
    # another payload's sequence, never the original payload (injective).
    bad = sym.copy()
    bad[0, 0] = (bad[0, 0] + 1) % 8
    j = int(np.flatnonzero(sym[1] != sym[1, 0])[0])
    bad[1, [0, j]] = bad[1, [j, 0]]
    back, valid = cpu.ccdm_dematch(bad, cc.composition, cc.k, cc._buffer_bits)
    assert not valid[0]
    assert not valid[1] or not np.array_equal(back[1], u[1])
    assert valid[2:].all()
    np.testing.assert_array_equal(back[2:], u[2:])


def test_oracle_dematcher_guards_its_buffer():
    cc = CCDM(AMPLITUDES, num_symbols=50)
    u = np.random.default_rng(3).integers(0, 2, size=(1, cc.k), dtype=np.int8)
    sym = cpu.ccdm_match(u, cc.composition)
    assert cpu.ccdm_dematch(sym, cc.composition, cc.k, cc._buffer_bits)[1][0]
    assert not cpu.ccdm_dematch(sym, cc.composition, cc.k, cc.k + 1)[1][0]


def test_rate_approaches_entropy():
    """Rate loss H - k/n is positive and O(log n / n)."""
    p = np.asarray(AMPLITUDES.pmf)
    entropy = float(-(p * np.log2(p)).sum())
    losses = []
    for n in (100, 1_000, 10_000, 100_000):
        loss = entropy - CCDM(AMPLITUDES, num_symbols=n).k / n
        assert 0 < loss < 4 * np.log2(n) / n
        losses.append(loss)
    assert np.all(np.diff(losses) < 0)


def test_validation():
    with pytest.raises(TypeError, match="Constellation"):
        CCDM("qam", num_symbols=10)
    with pytest.raises(ValueError, match="num_symbols"):
        CCDM(AMPLITUDES, num_symbols=0)
    with pytest.raises(ValueError, match="no information bit"):
        CCDM(AMPLITUDES, num_symbols=2)
    assert CCDM(AMPLITUDES, num_symbols=10) == CCDM(AMPLITUDES, num_symbols=10)
    assert "num_symbols=10" in repr(CCDM(AMPLITUDES, num_symbols=10))
    code = seeded_code("r1/2-z27")
    with pytest.raises(ValueError, match="below"):  # rate 1/2 < (m - 1) / m = 3/4
        PAS(code, Constellation.qam(256).shaped(nu=0.05))
    with pytest.raises(TypeError, match="code object"):
        PAS("ldpc", Constellation.qam(16))


def test_pas_layout():
    code = seeded_code("r5/6-z360")
    pas = PAS(code, Constellation.qam(256).shaped(nu=0.05))
    m = 4
    dims = code.n // m
    assert pas.n == code.n and pas.matcher.num_symbols == dims
    assert pas.k == pas.matcher.k + code.k - (m - 1) * dims
    # Per real dimension the sign, then its m - 1 amplitude bits (the first
    # (m - 1) dims information positions, in order).
    perm = pas.permutation.reshape(dims, m)
    np.testing.assert_array_equal(perm[:, 1:].ravel(), np.arange((m - 1) * dims))
    assert set(perm[:, 0]) == set(range((m - 1) * dims, code.n))
    # The amplitude alphabet: positive levels, labels without the sign bit,
    # twice their prior.
    amps = pas.matcher.constellation
    assert amps.order == 8 and np.all(np.asarray(amps.points) > 0)
    assert np.isclose(np.asarray(amps.pmf).sum(), 1.0)
