"""Serving engine host logic (CPU): admission against ring capacity, ring
wrap-around, launch grouping and completion order, through the C ABI in a
dry run (no device work), and the arrival schedule and driver with a fake
clock."""

import numpy as np
import pytest

from commkit_amd import MinSum
from commkit_amd.serve import Server, drive, poisson_schedule
from commkit_amd.workloads import seeded_code


@pytest.fixture(scope="module")
def codes():
    return [seeded_code("r1/2-z27"), seeded_code("r2/3-z96")]


def test_admission_and_ring_reuse(codes):
    s = Server(codes, MinSum(), dry_run=True, capacities=[10, 4], max_inflight=2)
    a = s.admit(0, 6, 1)
    assert a is not None and a.first == 0
    b = s.admit(0, 4, 2)
    assert b is not None and b.first == 6
    assert s.admit(0, 1, 3) is None  # ring of code 0 full
    assert s.usage(0)[:2] == (10, 10)
    c = s.admit(1, 4, 3)
    assert c is not None and c.first == 0
    with pytest.raises(RuntimeError):
        s.admit(1, 5, 4)  # larger than the ring: never fits
    with pytest.raises(RuntimeError):
        s.admit(0, 1, 1)  # tag in use
    s.submit(1)
    s.submit(3)
    assert s.step() == 10  # request 2 not submitted: not launched
    with pytest.raises(RuntimeError):
        s.release(1)  # not polled as completed yet
    assert sorted(s.poll()) == [1, 3]
    s.release(1)
    assert s.usage(0)[0] == 4  # request 1's frames are free again
    # Slots 0..5 are free again, 6..9 held by request 2.
    assert s.admit(0, 7, 5) is None
    d = s.admit(0, 3, 5)
    assert d is not None and d.first == 0  # wraps to the ring start
    s.submit(2)
    s.submit(5)
    assert s.step() == 7
    assert sorted(s.poll()) == [2, 5]
    s.release(5)
    assert s.usage(0)[0] == 7  # request 2 still held: released in admission order
    s.release(2)
    s.release(3)
    assert s.usage(0)[0] == 0 and s.usage(1)[0] == 0
    # A request that does not fit before the ring's end skips to its start.
    e = s.admit(0, 5, 6)
    assert e is not None and e.first == 3
    s.submit(6)
    s.step()
    s.poll()
    s.release(6)
    f = s.admit(0, 5, 7)  # slots 8, 9 skipped
    assert f is not None and f.first == 0
    assert s.usage(0)[0] == 7
    s.close()


def test_contiguous_requests_share_a_class_and_inflight_limit(codes):
    s = Server(codes, MinSum(), dry_run=True, capacities=[64, 64], max_inflight=1)
    for t in range(5):
        assert s.admit(t % 2, 3, t) is not None
        s.submit(t)
    assert s.step() == 15
    assert s.usage(0)[2] == 1
    for t in range(5, 7):
        s.admit(0, 2, t)
        s.submit(t)
    # The single workspace is busy until polled; step reaps completions first.
    assert s.step() == 4
    done = s.poll()
    assert sorted(done) == list(range(7))


def test_schedule_and_driver_with_fake_clock(codes):
    sched = poisson_schedule(2000.0, 0.5, streams=8, num_codes=2, max_frames=16, seed=3)
    assert len(sched) > 500 and np.all(np.diff(sched.t) >= 0)
    assert set(np.unique(sched.code)) == {0, 1}
    assert sched.frames.min() >= 1 and sched.frames.max() <= 16
    assert sched.snr_bin.min() >= 0 and sched.snr_bin.max() < 16
    # Consecutive requests of one stream move one Es/N0 bin.
    for s in range(8):
        b = sched.snr_bin[sched.stream == s]
        assert np.all(np.abs(np.diff(b)) == 1)

    t = [0.0]

    def clock():
        t[0] += 2e-4
        return t[0]

    srv = Server(codes, MinSum(), dry_run=True, capacities=[40, 40], max_inflight=2)
    filled = []
    res = drive(srv, sched, lambda slot, i: filled.append(i), clock=clock)
    assert res.requests == len(sched) and sorted(filled) == list(range(len(sched)))
    assert res.latency_ms.size == res.frames and np.all(res.latency_ms >= 0)
    assert res.rejected > 0  # 40-frame rings: bursts wait for room
    assert srv.usage(0)[0] == 0 and srv.usage(1)[0] == 0


def test_launch_frame_cap_takes_oldest_whole_requests(codes):
    """With a frame cap a launch takes whole requests in submission order
    until the next one would exceed the cap (always at least one), so a
    burst is launched in pieces; a request larger than the cap goes alone."""
    s = Server(codes, MinSum(), dry_run=True, capacities=[64, 64], max_inflight=8,
               max_launch_frames=10)  # fmt: skip
    sizes = [(0, 4), (1, 5), (0, 3), (1, 12), (0, 2)]
    for t, (code, f) in enumerate(sizes):
        assert s.admit(code, f, t) is not None
    for t in (1, 0, 2, 3, 4):  # submission order differs from admission order
        s.submit(t)
    assert s.step() == 9  # requests 1 and 0; request 2 would make 12
    assert s.step() == 3  # request 2; request 3 would make 15
    assert s.step() == 12  # request 3 alone, above the cap
    assert s.step() == 2
    assert s.step() == 0
    assert sorted(s.poll()) == [0, 1, 2, 3, 4]
    assert s.max_launch_frames == 10
    s.close()
