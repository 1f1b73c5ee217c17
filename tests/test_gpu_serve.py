"""Serving engine on the GPU: requests of three codes served in a mix of
traffic (admitted, submitted and launched in waves, several launches in
flight, with and without a frame cap per launch) decode exactly as each
request decoded alone, for min-sum and for sum-product; requests that carry
received symbols, demapped into their ring slots, give the LLRs and
decoded outputs of compute_llr followed by decode."""

import numpy as np
import pytest

from commkit_amd import MinSum, SumProduct
from commkit_amd.coding.ldpc import _decode
from commkit_amd.serve import Server
from commkit_amd.workloads import seeded_code

pytestmark = pytest.mark.gpu


def _serve(srv, reqs, fill, check, per_turn=3, before_step=None):
    """Admit up to ``per_turn`` requests a turn (as room allows), step, poll
    and check every completed request; ``before_step()`` runs after a turn's
    fills, before its step."""
    pending = list(reqs)
    live = {}
    checked = 0
    turns = 0
    while checked < len(reqs):
        turns += 1
        assert turns < 10_000
        for _ in range(per_turn):
            if not pending:
                break
            req = pending[0]
            slot = srv.admit(req[1], req[2], req[0])
            if slot is None:
                break
            fill(slot, req)
            srv.submit(req[0])
            live[req[0]] = req
            pending.pop(0)
        if before_step is not None:
            before_step()
        srv.step()
        for tag in srv.poll():
            check(srv.slot(tag), live.pop(tag))
            srv.release(tag)
            checked += 1


@pytest.mark.parametrize("method", [MinSum(max_iterations=25), SumProduct(max_iterations=20)],
                         ids=["min-sum", "sum-product"])  # fmt: skip
@pytest.mark.parametrize("cap", [None, 16], ids=["uncapped", "cap16"])
def test_requests_identical_alone_and_in_the_mix(gpu, method, cap):
    import torch

    codes = [seeded_code("r5/6-z360"), seeded_code("r2/3-z96"), seeded_code("r1/2-z27")]
    rng = np.random.default_rng(7)
    srv = Server(codes, method, capacities=[40, 120, 200], max_inflight=3, max_launch_frames=cap)
    reqs = []
    for tag in range(36):
        c = int(rng.integers(0, 3))
        R = int(rng.integers(1, 24))
        mean = rng.uniform(0.5, 2.5, size=(R, 1))
        llr = (mean + 1.6 * rng.standard_normal((R, codes[c].n))).astype(np.float32)
        reqs.append((tag, c, R, torch.from_numpy(llr).cuda()))

    def fill(slot, req):
        slot.llrs.copy_(req[3])

    def check(slot, req):
        tag, c, _, llr = req
        bits, app, valid, iters = _decode(codes[c], llr, method)
        assert torch.equal(slot.bits, bits), f"request {tag}: bits"
        assert torch.equal(slot.app, app), f"request {tag}: APP"
        assert torch.equal(slot.iterations, iters), f"request {tag}: iterations"
        assert torch.equal(slot.valid.bool(), valid), f"request {tag}: valid"

    _serve(srv, reqs, fill, check)
    srv.close()


@pytest.mark.parametrize("batched", [False, True], ids=["per-request", "per-turn"])
def test_symbols_demapped_into_slots(gpu, batched):
    """Requests of received symbols at their own noise levels, demapped into
    the ring slot by a RowsDemapper call each or queued on a
    SegmentsDemapper and launched once a turn: the slot's LLRs equal
    compute_llr's and the decoded request equals decode() of those LLRs."""
    import torch

    from commkit_amd import Constellation, compute_llr
    from commkit_amd.mapping.demap import RowsDemapper, SegmentsDemapper
    from commkit_amd.operating import waterfall_db
    from commkit_amd.workloads import make_frames

    names = ("r2/3-z96", "r1/2-z27")
    codes = [seeded_code(n) for n in names]
    orders = (16, 64)
    method = MinSum(max_iterations=30)
    srv = Server(codes, method, capacities=[64, 64], max_inflight=2, max_launch_frames=24)
    kind = SegmentsDemapper if batched else RowsDemapper
    demap = {o: kind(Constellation.qam(o), "maxlog", gpu) for o in orders}
    rng = np.random.default_rng(3)
    reqs = []
    for tag in range(20):
        c = int(rng.integers(0, 2))
        order = orders[int(rng.integers(0, 2))]
        R = int(rng.integers(1, 12))
        es = waterfall_db(names[c], order) + float(rng.uniform(-0.5, 1.0))
        fr = make_frames(codes[c], Constellation.qam(order), es, R, seed=50 + tag, device=gpu)
        inv = torch.full((R,), 1.0 / float(fr.noise_var[0]), dtype=torch.float64, device=gpu)
        reqs.append((tag, c, R, order, fr.rx.contiguous(), inv, float(fr.noise_var[0])))

    def fill(slot, req):
        if batched:
            demap[req[3]].add(req[4], slot.llrs, inv_s2=req[5])
        else:
            demap[req[3]](req[4], req[5], slot.llrs)

    def flush():
        for d in demap.values():
            d.flush()

    def check(slot, req):
        tag, c, R, order, rx, _, nv = req
        llr = compute_llr(rx, noise_var=nv, constellation=Constellation.qam(order))
        assert torch.equal(slot.llrs, llr.reshape(R, codes[c].n)), f"request {tag}: LLRs"
        bits, app, valid, iters = _decode(codes[c], slot.llrs, method)
        assert torch.equal(slot.bits, bits), f"request {tag}: bits"
        assert torch.equal(slot.app, app), f"request {tag}: APP"
        assert torch.equal(slot.iterations, iters), f"request {tag}: iterations"

    _serve(srv, reqs, fill, check, per_turn=5 if batched else 2,
           before_step=flush if batched else None)  # fmt: skip
    srv.close()
