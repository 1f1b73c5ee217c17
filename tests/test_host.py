"""Host logic on the CPU: codes, schedules, encoder, constellations, argument checks."""

import numpy as np
import pytest
import scipy.sparse

from commkit_amd import Constellation, LDPCCode, MinSum, SumProduct, compute_llr, decode
from commkit_amd.coding.gf2 import inverse, rank, row_reduce
from commkit_amd.workloads import CODES, seeded_base, seeded_code

from .codes import SMALL_QC_BASE, TREE_ROWS, encode_host, h_of, rows_in_schedule, small_qc, syndrome
from .codes import quiet, tree_code


class TestCodes:
    def test_quasi_cyclic_circulants(self):
        code = small_qc()
        h = code.parity_check().toarray()
        assert h.shape == (20, 40)
        for (i, j), s in np.ndenumerate(SMALL_QC_BASE):
            block = h[5 * i : 5 * i + 5, 5 * j : 5 * j + 5]
            if s < 0:
                assert not block.any()
            else:
                np.testing.assert_array_equal(block, np.roll(np.eye(5), s, axis=1))

    def test_from_parity_check_forms_agree(self):
        h = h_of(TREE_ROWS, 8)
        a = LDPCCode.from_parity_check(h)
        b = LDPCCode.from_parity_check(scipy.sparse.csr_array(h))
        assert a == b and hash(a) == hash(b)
        assert (a.n, a.k, a.num_checks) == (8, 4, 4)
        assert repr(a) == "LDPCCode(n=8, k=4)"

    def test_validation(self):
        with pytest.raises(ValueError, match="only 0 and 1"):
            LDPCCode.from_parity_check([[2, 1, 0]])
        with pytest.raises(ValueError, match="at least two"):
            LDPCCode.from_parity_check(h_of([[0, 1], [2]], 4))
        with pytest.raises(ValueError, match="duplicate"):
            LDPCCode([0, 0, 1, 1], [0, 0, 1, 2], 2, 4)
        with pytest.raises(ValueError, match="out of range"):
            LDPCCode([0, 0], [0, 9], 1, 4)
        with pytest.raises(ValueError, match="num_checks"):
            LDPCCode([0, 0], [0, 1], 2, 2)
        with pytest.raises(TypeError, match="must be integers"):
            LDPCCode(np.array([0.0, 0.0]), np.array([0.1, 2.1]), 1, 4)
        with pytest.raises(TypeError, match="num_checks must be an integer"):
            LDPCCode([0, 0, 1, 1], [0, 2, 1, 3], 2.9, 4)
        with pytest.raises(ValueError, match="parity bits"):
            LDPCCode.from_parity_check(h_of(TREE_ROWS, 8), punctured=5)
        with pytest.raises(ValueError, match="lifting"):
            LDPCCode.quasi_cyclic(SMALL_QC_BASE, lifting=0)
        with pytest.raises(ValueError, match="shifts"):
            LDPCCode.quasi_cyclic(SMALL_QC_BASE - 2, lifting=5)

    def test_methods_validation(self):
        with pytest.raises(ValueError, match="scale"):
            MinSum(scale=0.0)
        with pytest.raises(ValueError, match="offset"):
            MinSum(offset=-1.0)
        with pytest.raises(ValueError, match="max_iterations"):
            SumProduct(max_iterations=0)
        with pytest.raises(TypeError, match="max_iterations"):
            MinSum(max_iterations=True)
        with pytest.raises(ValueError, match="schedule"):
            MinSum(schedule="zigzag")
        assert MinSum() == MinSum(0.875, 0.0, 50, True, "layered")

    def test_warns_on_4_cycles(self, caplog):
        import logging

        with caplog.at_level(logging.WARNING, logger="commkit_amd"):
            LDPCCode.from_parity_check(h_of([[0, 1, 2], [0, 1, 3]], 5))
        assert "4-cycles" in caplog.text


class TestSchedule:
    @pytest.mark.parametrize("name", ["small", "tree", "r1/2-z27", "r4/5-z360"])
    def test_first_fit_layers(self, name):
        """Layers share no bit; the rows are H's rows; and each row's layer is
        the first-fit colour: every lower colour is held by an earlier row
        that shares a bit with it, and its own colour by no such row."""
        code = {"small": small_qc, "tree": tree_code}.get(name, lambda: seeded_code(name))()
        row_ptr, cols, layer_ptr = code._schedule
        order = code._schedule_rows
        assert row_ptr[-1] == code.checks.size
        for layer in range(layer_ptr.size - 1):
            bits = cols[row_ptr[layer_ptr[layer]] : row_ptr[layer_ptr[layer + 1]]]
            assert np.unique(bits).size == bits.size
        got = sorted(tuple(r) for r in rows_in_schedule(code))
        h = code.parity_check()
        assert got == sorted(tuple(h.indices[h.indptr[i] : h.indptr[i + 1]].tolist())
                             for i in range(h.shape[0]))  # fmt: skip
        if code.num_checks > 400:
            return  # the brute-force colour check below is quadratic
        color = np.repeat(np.arange(layer_ptr.size - 1), np.diff(layer_ptr))
        row_color = np.empty(code.num_checks, int)
        row_color[order] = color
        hd = h.toarray().astype(bool)
        for r in range(code.num_checks):
            earlier = [q for q in range(r) if (hd[q] & hd[r]).any()]
            held = {row_color[q] for q in earlier}
            assert row_color[r] not in held
            assert all(c in held for c in range(row_color[r]))

    @pytest.mark.parametrize("name", list(CODES))
    def test_qc_layers_are_whole_block_rows(self, name):
        code = seeded_code(name)
        z = CODES[name].lifting
        order = code._schedule_rows
        blocks = order.reshape(-1, z)
        np.testing.assert_array_equal(blocks, blocks[:, :1] + np.arange(z))
        qc = code._qc_view
        assert qc is not None
        np.testing.assert_array_equal(qc["row_id"], blocks[:, 0] // z)
        # slots in ascending block column within each block row
        for br in range(qc["row_ptr"].size - 1):
            s = qc["slot_col"][qc["row_ptr"][br] : qc["row_ptr"][br + 1]]
            assert np.all(np.diff(s) > 0)

    def test_var_view_ascending_checks(self):
        code = seeded_code("r1/2-z27")
        row_ptr, cols, _ = code._schedule
        var_ptr, var_edges = code._var_view
        nat = np.repeat(code._schedule_rows, np.diff(row_ptr))
        for v in range(0, code.num_variables, 37):
            e = var_edges[var_ptr[v] : var_ptr[v + 1]]
            assert np.all(cols[e] == v) and np.all(np.diff(nat[e]) > 0)
        assert var_ptr[-1] == cols.size


class TestEncoder:
    @pytest.mark.parametrize("kind", ["inverse", "dense"])
    def test_codewords_satisfy_parity(self, kind):
        code = small_qc() if kind == "inverse" else tree_code()
        assert code._encoder.kind == kind
        u = np.random.default_rng(1).integers(0, 2, (6, code.k))
        c = encode_host(code, u)
        assert not syndrome(code, c).any()
        np.testing.assert_array_equal(c[:, code._encoder.info], u)

    def test_accumulate_kind(self):
        # dual-diagonal parity part: accumulator encoding
        m, k = 6, 6
        rows, cols = [], []
        rng = np.random.default_rng(2)
        for i in range(m):
            for j in rng.choice(k, 3, replace=False):
                rows.append(i)
                cols.append(j)
            rows.append(i)
            cols.append(k + i)
            if i:
                rows.append(i)
                cols.append(k + i - 1)
        code = LDPCCode(np.array(rows), np.array(cols), m, m + k)
        assert code._encoder.kind == "accumulate"
        u = rng.integers(0, 2, (5, k))
        c = encode_host(code, u)
        assert not syndrome(code, c).any()

    @pytest.mark.parametrize("name", ["r1/2-z27", "r2/3-z96"])
    def test_seeded_codes(self, name):
        spec = CODES[name]
        base = seeded_base(name)
        assert base.shape == (spec.base_rows, spec.base_cols)
        assert base.max() < spec.lifting
        rw = (base >= 0).sum(axis=1)
        cw = (base >= 0).sum(axis=0)
        assert rw.min() >= 2 and rw.max() <= 32
        if spec.rule == "weights":
            assert set(cw.tolist()) <= {3, 4}
        else:
            assert cw.min() >= 2
        code = seeded_code(name)
        assert code.k == code.n - code.num_checks
        u = np.random.default_rng(3).integers(0, 2, (3, code.k))
        assert not syndrome(code, encode_host(code, u)).any()

    @pytest.mark.parametrize("name", list(CODES))
    def test_row_weight_cap_rejects_no_draw(self, name):
        """The 32-edge row cap never decides a seeded code: no draw up to the
        accepted one has a row heavier than 32, so every code is the draw it
        would be without the cap."""
        from commkit_amd.workloads import _draw

        spec = CODES[name]
        rng = np.random.Generator(np.random.PCG64(spec.seed))
        accepted = seeded_base(name)
        for _ in range(10_000):
            base = _draw(spec, rng)
            assert (base >= 0).sum(axis=1).max() <= 32
            if np.array_equal(base, accepted):
                break
        else:
            pytest.fail("accepted base matrix not reached")

    def test_gf2(self):
        rng = np.random.default_rng(4)
        for _ in range(5):
            a = rng.integers(0, 2, (12, 12)).astype(np.int8)
            if rank(a) < 12:
                with pytest.raises(ValueError, match="singular"):
                    inverse(a)
                continue
            inv = inverse(a)
            np.testing.assert_array_equal((a.astype(int) @ inv) % 2, np.eye(12, dtype=int))
        h = rng.integers(0, 2, (5, 9)).astype(np.int8)
        rref, piv = row_reduce(h, reverse=True)
        assert np.all(np.diff(piv) < 0)
        np.testing.assert_array_equal(rref[:, piv], np.eye(piv.size, dtype=np.int8))
        # the pivot columns are the rightmost independent ones
        for j in range(9):
            right = [c for c in range(j + 1, 9) if c in piv]
            indep = rank(h[:, right + [j]]) > rank(h[:, right]) if right else h[:, j].any()
            assert (j in piv) == bool(indep)


class TestConstellation:
    @pytest.mark.parametrize("order", [4, 16, 64, 256])
    def test_gray_qam(self, order):
        c = Constellation.qam(order)
        assert abs(c.power() - 1.0) < 1e-12
        labels = c.bit_labels.astype(int)
        pts = c.points
        d = np.abs(pts[:, None] - pts[None, :])
        dmin = d[d > 0].min()
        near = np.isclose(d, dmin)
        diff = (labels[:, None, :] != labels[None, :, :]).sum(axis=2)
        assert np.all(diff[near] == 1)  # nearest neighbours differ in one bit
        assert c._axes is not None
        k = c.bits_per_symbol
        np.testing.assert_array_equal(c._axes[0].bits, np.arange(k // 2))
        np.testing.assert_array_equal(c._axes[1].bits, np.arange(k // 2, k))

    @pytest.mark.parametrize("order", [32, 128])
    def test_cross_qam_matches_the_reference(self, order):
        """Cross QAM: the point of every label equals the reference's
        (golden vectors written from commkit's own _gray_qam_cross by
        tests/golden/make_cross_qam.py), scaled to unit power; the points
        form the cross (the 2^ceil(k/2) x 2^floor(k/2) rectangle with an
        eighth of the columns folded on each side), and most nearest
        neighbours differ in one bit (quasi-Gray)."""
        import json
        import os

        with open(os.path.join(os.path.dirname(__file__), "golden", "cross_qam.json")) as f:
            golden = np.array([complex(a, b) for a, b in json.load(f)[str(order)]])
        c = Constellation.qam(order)
        assert abs(c.power() - 1.0) < 1e-12
        np.testing.assert_allclose(c.points, golden / np.sqrt(np.mean(np.abs(golden) ** 2)),
                                   rtol=0, atol=1e-15)  # fmt: skip
        # the cross: distinct odd-integer points within +-(W - 2b - 1) on both
        # axes, none in a corner where both |I| and |Q| exceed H
        k = int(np.log2(order))
        W, H = 1 << ((k + 1) // 2), 1 << (k // 2)
        side = W - 2 * (W // 8) - 1
        grid = {(int(p.real), int(p.imag)) for p in golden}
        assert len(grid) == order and all(i % 2 and q % 2 for i, q in grid)
        assert max(max(abs(i), abs(q)) for i, q in grid) == side
        assert not any(abs(i) > H and abs(q) > H for i, q in grid)
        assert c._axes is None  # not separable: the M-point demapper
        labels = c.bit_labels.astype(int)
        d = np.abs(c.points[:, None] - c.points[None, :])
        near = np.isclose(d, d[d > 0].min())
        diff = (labels[:, None, :] != labels[None, :, :]).sum(axis=2)
        assert (diff[near] == 1).mean() > 0.8

    def test_shaped_maxwell_boltzmann(self):
        c = Constellation.qam(256).shaped(nu=0.05)
        assert abs(c.pmf.sum() - 1) < 1e-12 and abs(c.power() - 1) < 1e-12
        # P ~ exp(-nu |s'|^2) on the odd-integer grid (min distance 2)
        s = Constellation.qam(256).points
        grid = s / (np.abs(s[0] - s[1]) / 2)
        want = np.exp(-0.05 * np.abs(grid) ** 2)
        np.testing.assert_allclose(c.pmf, want / want.sum(), rtol=1e-12)
        assert c._axes is not None and c._axes[0].pmf is not None
        assert Constellation.psk(8)._axes is None

    def test_entropy_target(self):
        c = Constellation.qam(64).shaped(entropy=5.0)
        assert abs(c.entropy - 5.0) < 1e-9


class TestArgumentChecks:
    def test_compute_llr_validation(self):
        c = Constellation.qam(16)
        x = np.zeros(4, np.complex64)
        with pytest.raises(ValueError, match="one of noise_var and esn0_db"):
            compute_llr(x, constellation=c)
        with pytest.raises(ValueError, match="one of noise_var and esn0_db"):
            compute_llr(x, constellation=c, noise_var=0.1, esn0_db=3.0)
        with pytest.raises(ValueError, match="Unknown method"):
            compute_llr(x, constellation=c, noise_var=0.1, method="bogus")
        with pytest.raises(ValueError, match="constellation is required"):
            compute_llr(x, noise_var=0.1)

    def test_decode_validation(self):
        code = small_qc()
        with pytest.raises(TypeError, match="code must be a code object"):
            decode(np.zeros(40, np.float32), code="not a code")
        with pytest.raises(TypeError, match="does not decode"):
            decode(np.zeros(40, np.float32), code=code, method=object())


def _rank_mod2(m: np.ndarray) -> int:
    """Rank over GF(2) by plain Gaussian elimination on Python ints (one int
    per row, an independent check of gf2.rank)."""
    rows = [int("".join(str(int(b)) for b in r), 2) for r in np.asarray(m) & 1] if m.size else []
    rank = 0
    while rows:
        pivot = max(rows)
        rows.remove(pivot)
        if pivot == 0:
            break
        top = pivot.bit_length() - 1
        rows = [r ^ pivot if (r >> top) & 1 else r for r in rows]
        rank += 1
    return rank


class TestPivotChoice:
    """``row_reduce`` pivots, which fix the dense encoder's information
    positions (``commkit/coding/_gf2.py:45-88``, ``ldpc.py:445-481``): a
    column is a pivot exactly when it is independent of the columns taken
    before it, in ascending order (forward) or from the last column back
    (reverse) -- the greedy basis of the column matroid, checked here against
    ranks of column ranges."""

    @pytest.mark.parametrize("seed", range(6))
    def test_pivots_are_the_greedy_basis(self, seed):
        rng = np.random.default_rng(seed)
        r, n = [(4, 9), (6, 6), (7, 12), (5, 16), (9, 14), (3, 20)][seed]
        m = (rng.random((r, n)) < 0.35).astype(np.int8)
        if seed == 1:
            m[2] = m[0] ^ m[1]  # a dependent row
        _, fwd = row_reduce(m)
        _, rev = row_reduce(m, reverse=True)
        want_fwd = [c for c in range(n) if _rank_mod2(m[:, : c + 1]) > _rank_mod2(m[:, :c])]
        want_rev = [c for c in range(n - 1, -1, -1)
                    if _rank_mod2(m[:, c:]) > _rank_mod2(m[:, c + 1 :])]  # fmt: skip
        assert fwd.tolist() == want_fwd
        assert rev.tolist() == want_rev
        assert len(want_fwd) == len(want_rev) == _rank_mod2(m) == rank(m)

    def test_dense_encoder_information_positions(self):
        code = tree_code()
        assert code._encoder.kind == "dense"
        h = code.parity_check().toarray().astype(np.int8)
        n = h.shape[1]
        piv = [c for c in range(n - 1, -1, -1) if _rank_mod2(h[:, c:]) > _rank_mod2(h[:, c + 1 :])]
        np.testing.assert_array_equal(code._encoder.info, np.setdiff1d(np.arange(n), piv))



def test_tuned_tiles_are_keyed_by_the_code():
    """A tuned tile applies to the exact code it was tuned for: another code
    of the same length and lifting (here a rate-2/3 base graph at N = 21600,
    Z = 360, beside the tuned rate-5/6 code) takes the default tile."""
    from commkit_amd.coding import ldpc

    tuned = seeded_code("r5/6-z360")
    assert ldpc.tuned_tiles(tuned, 0, 0) is not None
    rng = np.random.Generator(np.random.PCG64(7))
    mb, nb = 20, 60
    base = np.where(rng.random((mb, nb)) < 0.1, rng.integers(0, 360, (mb, nb)), -1)
    base[:, nb - mb:] = -1
    # parity part block lower bidiagonal with identities on the diagonal: invertible
    base[np.arange(mb), nb - mb + np.arange(mb)] = 0
    base[np.arange(1, mb), nb - mb + np.arange(mb - 1)] = rng.integers(0, 360, mb - 1)
    base[0, : nb - mb] = np.where(base[0, : nb - mb] < 0, 5, base[0, : nb - mb])
    other = quiet(LDPCCode.quasi_cyclic, base, lifting=360)
    assert other.num_variables == tuned.num_variables
    assert ldpc.code_fingerprint(other) != ldpc.code_fingerprint(tuned)
    assert ldpc.tuned_tiles(other, 0, 0) is None
    ms = ldpc.method_struct(other, MinSum())
    assert (ms.kernel, ms.frames_per_group, ms.lanes_per_row, ms.state_in_lds) == (0, 0, 0, -1)
    # the fingerprint follows the matrix, not the object
    again = LDPCCode.quasi_cyclic(seeded_base("r5/6-z360"), lifting=360)
    assert ldpc.code_fingerprint(again) == ldpc.code_fingerprint(tuned)
