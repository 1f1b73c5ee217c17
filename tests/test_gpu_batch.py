"""Persistent batched decoder against the per-call decoder (GPU).

Every frame of a batch must decode exactly as the per-call decoder decodes
it alone (which the per-call tests hold bit-exact to the CPU oracle for
min-sum): bits, APP LLRs, iteration counts and valid flags, in direct
launches and in replays of the captured HIP graph, with other GPU work
between the replays.  The mixed workload's batched launch is also held to
the CPU oracle directly.
"""

import numpy as np
import pytest

from commkit_amd import MinSum, SumProduct
from commkit_amd.batch import BatchDecoder
from commkit_amd.coding.ldpc import _decode
from commkit_amd.workloads import seeded_code

from .codes import encode_host
from .test_gpu_ldpc import oracle_decode

pytestmark = pytest.mark.gpu

NAMES = ["r5/6-z360", "r2/3-z96", "r1/2-z27"]
FRAMES = [5, 24, 70]


def _llrs(code, rows, seed):
    r = np.random.default_rng(seed)
    mean = r.uniform(0.6, 2.5, size=(rows, 1))  # uneven convergence per frame
    return (mean + 1.5 * r.standard_normal((rows, code.n))).astype(np.float32)


def _alone(code, x, method):
    # The batched decoder runs two lanes per check row when two rows' lanes
    # fit its 768-lane workgroup, else one; for sum-product the per-call
    # reference must add a row's phi terms in the same order (same P).
    lanes = 2 if 2 * code._qc_view["lifting"] <= 768 else 1
    tiles = (2, 0, lanes, 0) if isinstance(method, SumProduct) else None
    bits, app, valid, iters = _decode(code, x, method, tiles=tiles)
    return [t.cpu().numpy() for t in (bits, app, valid.to(dtype=bits.dtype), iters)]


def _assert_class(res, want, label):
    got = [res.bits.cpu().numpy(), res.llrs.cpu().numpy(), res.valid.cpu().numpy().astype(np.int8),
           res.iterations.cpu().numpy()]  # fmt: skip
    for name, g, w in zip(("bits", "llrs", "valid", "iterations"), got, want):
        if not np.array_equal(g, w.astype(g.dtype)):
            bad = np.flatnonzero((g != w).reshape(g.shape[0], -1).any(axis=1))
            diff = float(np.abs(g.astype(np.float64) - w.astype(np.float64)).max())
            pytest.fail(f"{label}: {name} differ on frames {bad[:10].tolist()} (max {diff:.3g})")


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("method", [MinSum(scale=0.75, max_iterations=20),
                                    SumProduct(max_iterations=12)], ids=["min-sum", "sum-product"])
def test_batch_equals_per_call(gpu, graph, method):
    import torch

    codes = [seeded_code(n) for n in NAMES]
    bd = BatchDecoder(codes, FRAMES, method, graph=graph)
    for rep in range(3):
        inputs = [torch.from_numpy(_llrs(c, R, seed=10 * rep + i)).cuda()
                  for i, (c, R) in enumerate(zip(codes, FRAMES))]  # fmt: skip
        for dst, src in zip(bd.inputs, inputs):
            dst.copy_(src)
        bd.run()
        torch.cuda.synchronize()
        _ = (torch.randn(1000, device="cuda") * 2).sum().item()  # eager work between replays
        for i, (c, x) in enumerate(zip(codes, inputs)):
            _assert_class(bd.outputs[i], _alone(c, x, method), f"rep {rep} class {NAMES[i]}")
    print("frames per group", bd.frames_per_group, "grid", bd.plan.grid, "lds", bd.plan.lds_bytes)
    bd.close()


def test_batch_early_exit_and_empty_class(gpu):
    import torch

    codes = [seeded_code("r2/3-z96"), seeded_code("r1/2-z27")]
    u = np.random.default_rng(3).integers(0, 2, (6, codes[0].k))
    clean = ((1.0 - 2.0 * encode_host(codes[0], u)) * 3.0).astype(np.float32)
    bd = BatchDecoder(codes, [6, 0], MinSum(max_iterations=10))
    bd.inputs[0].copy_(torch.from_numpy(clean))
    bd.run()
    res = bd.outputs[0]
    assert bool(res.valid.all()) and int(res.iterations.max()) == 0
    np.testing.assert_array_equal(res.bits.cpu().numpy(), u)
    assert bd.outputs[1].bits.shape == (0, codes[1].k)
    bd.close()


def test_mixed_workload_batch_equals_per_code(gpu):
    """The mixed workload (three codes, four QAM orders, per-frame Es/N0),
    demapped per group into the code classes and decoded in one batched
    launch with the mixed benchmark's method (layered NMS 0.875, <= 50
    iterations), equals decoding each code's frames alone, and equals the
    CPU oracle decoding the same LLRs on every frame."""
    from commkit_amd import Constellation
    from commkit_amd.mapping.demap import llr_rows
    from commkit_amd.workloads import MIXED_ORDERS, make_mixed

    mx = make_mixed(240, seed=5)
    assert sum(mx.rows) == 240 and len(mx.groups) == 12
    method = MinSum(max_iterations=50)
    bd = BatchDecoder(mx.codes, mx.rows, method)
    consts = {o: Constellation.qam(o) for o in MIXED_ORDERS}
    for g in mx.groups:
        llr_rows(consts[g.order], g.frames.rx, g.frames.noise_var, "maxlog",
                 out=bd.inputs[g.code_index][g.row0:g.row1])  # fmt: skip
    bd.run()
    for i, c in enumerate(mx.codes):
        _assert_class(bd.outputs[i], _alone(c, bd.inputs[i], method), f"code {i}")
        want = oracle_decode(c, bd.inputs[i].cpu().numpy(), method)
        _assert_class(bd.outputs[i], [want[0], want[1], want[2].astype(np.int8), want[3]],
                      f"code {i} vs oracle")
    ok = sum(int(bd.outputs[g.code_index].valid[g.row0:g.row1].sum()) for g in mx.groups)
    assert ok > 120  # most frames sit above their waterfall
    bd.close()
