"""Serving engine on the GPU: requests of three codes served in a mix of
traffic (admitted, submitted and launched in waves, several launches in
flight) decode exactly as each request decoded alone."""

import numpy as np
import pytest

from commkit_amd import MinSum
from commkit_amd.coding.ldpc import _decode
from commkit_amd.serve import Server
from commkit_amd.workloads import seeded_code

pytestmark = pytest.mark.gpu


def test_requests_identical_alone_and_in_the_mix(gpu):
    import torch

    codes = [seeded_code("r5/6-z360"), seeded_code("r2/3-z96"), seeded_code("r1/2-z27")]
    method = MinSum(max_iterations=25)
    rng = np.random.default_rng(7)
    srv = Server(codes, method, capacities=[40, 120, 200], max_inflight=2)
    reqs = []
    for tag in range(36):
        c = int(rng.integers(0, 3))
        R = int(rng.integers(1, 24))
        mean = rng.uniform(0.5, 2.5, size=(R, 1))
        llr = (mean + 1.6 * rng.standard_normal((R, codes[c].n))).astype(np.float32)
        reqs.append((tag, c, torch.from_numpy(llr).cuda()))
    pending = list(reqs)
    live = {}
    checked = 0
    turns = 0
    while checked < len(reqs):
        turns += 1
        assert turns < 10_000
        # Admit a few per turn (waves of traffic), as room allows.
        for _ in range(3):
            if not pending:
                break
            tag, c, llr = pending[0]
            slot = srv.admit(c, llr.shape[0], tag)
            if slot is None:
                break
            slot.llrs.copy_(llr)
            srv.submit(tag)
            live[tag] = (c, llr)
            pending.pop(0)
        srv.step()
        for tag in srv.poll():
            c, llr = live.pop(tag)
            s = srv.slot(tag)
            bits, app, valid, iters = _decode(codes[c], llr, method)
            assert torch.equal(s.bits, bits), f"request {tag}: bits"
            assert torch.equal(s.app, app), f"request {tag}: APP"
            assert torch.equal(s.iterations, iters), f"request {tag}: iterations"
            assert torch.equal(s.valid.bool(), valid), f"request {tag}: valid"
            srv.release(tag)
            checked += 1
    srv.close()
