"""ck_ldpc_decode_static, the decoder entry point that allocates nothing
(GPU): through the C ABI, two calls in a row -- a large sum-product batch
whose workspace exceeds the 64 MiB static scratch (decoded in chunks), then
a smaller min-sum batch that reuses the scratch the first call left -- each
equal to ck_ldpc_decode with a caller-allocated workspace, bit for bit."""

import ctypes

import numpy as np
import pytest

from commkit_amd import _native
from commkit_amd import MinSum, SumProduct
from commkit_amd.coding.ldpc import _decode, device_code, method_struct
from commkit_amd.workloads import seeded_code

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _static(code, method, llrs):
    import torch

    R = int(llrs.shape[0])
    app = torch.full((R, code.n), float("nan"), dtype=torch.float32, device=llrs.device)
    bits = torch.full((R, code.k), -1, dtype=torch.int8, device=llrs.device)
    iters = torch.full((R,), -1, dtype=torch.int32, device=llrs.device)
    valid = torch.full((R,), 7, dtype=torch.uint8, device=llrs.device)
    dc = device_code(code, llrs.device)
    ms = method_struct(code, method)
    lib = _native.lib()
    _native.check(lib.ck_ldpc_decode_static(
        ctypes.byref(dc.struct), ctypes.byref(ms), llrs.data_ptr(), app.data_ptr(),
        bits.data_ptr(), iters.data_ptr(), valid.data_ptr(), R, _native.stream_of(llrs.device)))
    torch.cuda.synchronize()
    return bits, app, valid.bool(), iters


def _llrs(code, rows, mean, seed, dev):
    import torch

    r = np.random.default_rng(seed)
    x = (mean + 1.5 * r.standard_normal((rows, code.n))).astype(np.float32)
    return torch.from_numpy(x).to(dev)


def test_static_scratch_equals_workspace_decode(gpu):
    import torch

    sp_code = seeded_code("r5/6-z360")
    sp = SumProduct(max_iterations=6)
    per_frame = sp_code.checks.size * 4  # float32 messages per edge
    big = SCRATCH // per_frame + 37  # one more chunk than the scratch holds
    x1 = _llrs(sp_code, big, 1.0, 1, gpu)
    ms_code = seeded_code("r4/5-z360")
    ms = MinSum(max_iterations=12)
    x2 = _llrs(ms_code, 40, 1.4, 2, gpu)
    for code, method, x in ((sp_code, sp, x1), (ms_code, ms, x2)):
        got = _static(code, method, x)
        want = _decode(code, x, method)
        torch.cuda.synchronize()
        for name, g, w in zip(("bits", "app", "valid", "iterations"), got, want):
            assert torch.equal(g, w), f"{type(method).__name__}: {name} differ"
