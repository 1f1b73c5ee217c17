"""Serving: continuous batching of decode requests on one GPU.

:class:`Server` wraps the ``ck_serve_*`` engine of ``libck_hip.so``.  A
request is what a caller of :func:`commkit_amd.decode` would pass: ``R``
frames of one code.  The server owns one ring of frame buffers per code in
device memory, sized from a fixed HBM budget; :meth:`Server.admit` reserves
room for a request (``None`` when the budget is full: admission control),
the caller writes its LLRs into the returned slot and :meth:`Server.submit`
s it; :meth:`Server.step` launches every submitted request not yet launched
in one persistent batched decode (so requests join the next launch as soon
as they arrive), :meth:`Server.poll` returns the requests whose launch has
finished, and :meth:`Server.release` hands their room back.  The rings, the
launch workspaces and the engine's descriptors stay resident between
launches.

A frame decodes to the same bits, APP LLRs and iteration count whatever it
is batched with, so a request's result equals decoding it alone with
:func:`commkit_amd.decode` (the tests check this in a mix of traffic).

:func:`drive` runs an arrival schedule (:func:`poisson_schedule`: Poisson
request streams, sizes uniform in 1..256 frames, a per-stream Es/N0 random
walk) against a server and measures every frame's latency: completion seen
by the host minus arrival, on the host's monotonic clock.

No counterpart in the reference, which decodes one call at a time
(``commkit/coding/base.py:317-379``).
"""

from __future__ import annotations

import ctypes
import gc
import time
from dataclasses import dataclass, field
from typing import Any, Callable, Sequence

import numpy as np

from . import _native
from .coding.ldpc import LDPCCode, MinSum, SumProduct, device_code, method_struct

__all__ = ["Server", "Slot", "poisson_schedule", "burst_schedule", "Schedule", "drive",
           "LoadResult"]

CK_EBUSY = 4


class CkServeRing(ctypes.Structure):
    _fields_ = [("llrs", ctypes.c_void_p), ("app", ctypes.c_void_p), ("bits", ctypes.c_void_p),
                ("iterations", ctypes.c_void_p), ("valid", ctypes.c_void_p),
                ("capacity", ctypes.c_int64)]


class CkServeConfig(ctypes.Structure):
    _fields_ = [("codes", ctypes.c_void_p), ("rings", ctypes.c_void_p),
                ("num_codes", ctypes.c_int32), ("max_inflight", ctypes.c_int32),
                ("method", ctypes.c_void_p), ("dry_run", ctypes.c_int32),
                ("max_launch_frames", ctypes.c_int64)]


class CkServeSlot(ctypes.Structure):
    _fields_ = [("llrs", ctypes.c_void_p), ("app", ctypes.c_void_p), ("bits", ctypes.c_void_p),
                ("iterations", ctypes.c_void_p), ("valid", ctypes.c_void_p),
                ("first", ctypes.c_int64)]


class Slot:
    """An admitted request: its code and frame range in that code's ring,
    the device address of its LLR rows (``llrs_ptr``, 0 in a dry run), and
    (not in a dry run) device tensor views of its buffers, ``llrs``,
    ``app``, ``bits``, ``iterations`` and ``valid``, made on first use (a
    request loop that fills the LLRs through ``llrs_ptr`` and reads results
    with device work of its own never pays for the views)."""

    __slots__ = ("tag", "code", "first", "frames", "llrs_ptr", "_ring", "_views")
    _NAMES = ("llrs", "app", "bits", "iterations", "valid")

    def __init__(self, tag: int, code: int, first: int, frames: int, llrs_ptr: int = 0,
                 ring: dict[str, Any] | None = None) -> None:
        self.tag, self.code, self.first, self.frames = tag, code, first, frames
        self.llrs_ptr = llrs_ptr
        self._ring = ring
        self._views: dict[str, Any] = {}

    def _view(self, name: str) -> Any:
        if self._ring is None:
            return None
        v = self._views.get(name)
        if v is None:
            v = self._ring[name][self.first : self.first + self.frames]
            self._views[name] = v
        return v

    llrs = property(lambda self: self._view("llrs"))
    app = property(lambda self: self._view("app"))
    bits = property(lambda self: self._view("bits"))
    iterations = property(lambda self: self._view("iterations"))
    valid = property(lambda self: self._view("valid"))

    def __repr__(self) -> str:
        return (f"Slot(tag={self.tag}, code={self.code}, first={self.first}, "
                f"frames={self.frames})")


def frame_bytes(code: LDPCCode) -> int:
    """Device bytes a ring slot of ``code`` takes (LLRs in, APP, bits,
    iterations, valid out)."""
    return 8 * code.n + code.k + 5


class Server:
    """Continuous-batching decode server on one GPU.

    Parameters
    ----------
    codes : sequence of LDPCCode
        The codes requests may use (quasi-cyclic, at most 8).
    method : MinSum or SumProduct
        Layered decoding method of every request.
    hbm_budget : int
        Device bytes for the rings and launch workspaces; the rest of the
        budget after the workspaces is split evenly across the codes.
    max_inflight : int
        Launches in flight at once (each with its own workspace).
    device : torch device, optional
        Default: the current GPU.  Launches go on torch's current stream of
        that device at construction.
    dry_run : bool
        No device memory or work: the admission, ring and ordering logic
        only (for tests without a GPU).
    capacities : sequence of int, optional
        Frames per ring instead of the budget split.
    max_launch_frames : int, optional
        Frames one launch takes at most (whole requests, oldest first, at
        least one request).  Bounds how long a burst holds the device before
        a later request's launch; ``None`` or 0: no cap.
    """

    def __init__(self, codes: Sequence[LDPCCode], method: MinSum | SumProduct | None = None, *,
                 hbm_budget: int = 16 << 30, max_inflight: int = 2, device: Any = None,
                 dry_run: bool = False, capacities: Sequence[int] | None = None,
                 max_launch_frames: int | None = None) -> None:
        method = method or MinSum()
        if method.schedule != "layered":
            raise ValueError("serving runs the layered schedule")
        if not 1 <= len(codes) <= 8:
            raise ValueError("need 1 to 8 codes")
        self.codes = list(codes)
        self.method = method
        self.dry_run = dry_run
        self._lib = _native.lib()
        self._keep: list[Any] = []
        self._slots: dict[int, Slot] = {}
        import torch

        self.device = None
        if not dry_run:
            self.device = (torch.device("cuda", torch.cuda.current_device()) if device is None
                           else torch.device(device))  # fmt: skip
        dev_codes = []
        for c in self.codes:
            if c._qc_view is None:
                raise ValueError(f"{c!r} has no quasi-cyclic view")
            # A dry run's tables stay on the host: the engine never reads them.
            dc = device_code(c, torch.device("cpu") if dry_run else self.device)
            self._keep.append(dc)
            dev_codes.append(dc)
        self._ms = method_struct(self.codes[0], method)
        code_ptrs = (ctypes.c_void_p * len(codes))(*[ctypes.addressof(d.struct) for d in dev_codes])
        rings = (CkServeRing * len(codes))()
        self.max_launch_frames = int(max_launch_frames or 0)
        cfg = CkServeConfig(ctypes.addressof(code_ptrs), ctypes.addressof(rings), len(codes),
                            max_inflight, ctypes.addressof(self._ms), int(dry_run),
                            self.max_launch_frames)  # fmt: skip
        self._cfg, self._code_ptrs, self._rings = cfg, code_ptrs, rings
        # Workspace size first (rings only need capacities for it).
        for r in rings:
            r.capacity = 1
        ws = self._lib.ck_serve_workspace_bytes(ctypes.byref(cfg))
        if ws < 0:
            _native.check(self._lib.ck_serve_create(ctypes.byref(cfg), None, None,
                                                    ctypes.byref(ctypes.c_void_p())))  # fmt: skip
            raise RuntimeError("ck_serve_workspace_bytes failed")
        self.workspace_bytes = int(ws)
        if capacities is None:
            share = (int(hbm_budget) - max_inflight * self.workspace_bytes) // len(codes)
            capacities = [share // frame_bytes(c) for c in self.codes]
        if min(capacities) < 1:
            raise ValueError(f"HBM budget of {hbm_budget} bytes leaves no room for a frame")
        self.capacities = [int(x) for x in capacities]
        self.ring_tensors: list[dict[str, Any]] = []
        ws_ptrs = (ctypes.c_void_p * max_inflight)()
        if dry_run:
            for r, cap in zip(rings, self.capacities):
                r.capacity = cap
        else:
            with torch.cuda.device(self.device):
                for r, c, cap in zip(rings, self.codes, self.capacities):
                    t = {
                        "llrs": torch.zeros((cap, c.n), dtype=torch.float32, device=self.device),
                        "app": torch.empty((cap, c.n), dtype=torch.float32, device=self.device),
                        "bits": torch.empty((cap, c.k), dtype=torch.int8, device=self.device),
                        "iterations": torch.empty(cap, dtype=torch.int32, device=self.device),
                        "valid": torch.empty(cap, dtype=torch.uint8, device=self.device),
                    }
                    self.ring_tensors.append(t)
                    r.llrs, r.app, r.bits = t["llrs"].data_ptr(), t["app"].data_ptr(), t["bits"].data_ptr()
                    r.iterations, r.valid = t["iterations"].data_ptr(), t["valid"].data_ptr()
                    r.capacity = cap
                self._ws = [torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
                            for _ in range(max_inflight)]  # fmt: skip
                for i, w in enumerate(self._ws):
                    ws_ptrs[i] = w.data_ptr()
        stream = None if dry_run else _native.stream_of(self.device)
        self._h = ctypes.c_void_p()
        _native.check(self._lib.ck_serve_create(ctypes.byref(cfg), ws_ptrs, stream,
                                                ctypes.byref(self._h)))  # fmt: skip
        self._tags = (ctypes.c_uint64 * 4096)()

    # -- requests ---------------------------------------------------------

    def admit(self, code: int, frames: int, tag: int) -> Slot | None:
        """Room for ``frames`` frames of ``codes[code]``, or ``None`` when the
        budget is full now (retry after releasing completed requests)."""
        where = CkServeSlot()
        st = self._lib.ck_serve_admit(self._h, int(code), int(frames), int(tag), ctypes.byref(where))
        if st == CK_EBUSY:
            return None
        _native.check(st)
        s = Slot(int(tag), int(code), int(where.first), int(frames),
                 0 if self.dry_run else int(where.llrs or 0),
                 None if self.dry_run else self.ring_tensors[code])  # fmt: skip
        self._slots[s.tag] = s
        return s

    def submit(self, tag: int) -> None:
        _native.check(self._lib.ck_serve_submit(self._h, int(tag)))

    def step(self) -> int:
        """Launch the submitted requests not yet launched (oldest first, up to
        the frame cap); frames launched."""
        n = ctypes.c_int64()
        _native.check(self._lib.ck_serve_step(self._h, ctypes.byref(n)))
        return int(n.value)

    def poll(self) -> list[int]:
        """Tags of the requests completed since the last poll."""
        out: list[int] = []
        while True:
            cnt = ctypes.c_int32()
            _native.check(self._lib.ck_serve_poll(self._h, self._tags, len(self._tags),
                                                  ctypes.byref(cnt)))  # fmt: skip
            out.extend(self._tags[i] for i in range(cnt.value))
            if cnt.value < len(self._tags):
                return out

    def release(self, tag: int) -> None:
        _native.check(self._lib.ck_serve_release(self._h, int(tag)))
        self._slots.pop(int(tag), None)

    def slot(self, tag: int) -> Slot:
        return self._slots[int(tag)]

    def usage(self, code: int) -> tuple[int, int, int]:
        """(frames in use, capacity) of a code's ring, launches in flight."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _native.check(self._lib.ck_serve_usage(self._h, int(code), ctypes.byref(a),
                                               ctypes.byref(b), ctypes.byref(c)))  # fmt: skip
        return int(a.value), int(b.value), int(c.value)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h:
            _native.check(self._lib.ck_serve_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


# -- arrival schedules and the driver ----------------------------------------


@dataclass
class Schedule:
    """Requests in arrival order: time (s from start), stream, code index,
    frames, Es/N0 offset bin of the stream's random walk."""

    t: np.ndarray
    stream: np.ndarray
    code: np.ndarray
    frames: np.ndarray
    snr_bin: np.ndarray

    def __len__(self) -> int:
        return int(self.t.size)


def poisson_schedule(rate_per_s: float, duration_s: float, *, streams: int = 32,
                     num_codes: int = 3, max_frames: int = 256, snr_bins: int = 16,
                     seed: int = 0) -> Schedule:
    """``streams`` Poisson request streams of total rate ``rate_per_s`` over
    ``duration_s``: stream ``s`` sends code ``s % num_codes``, sizes uniform
    in ``1..max_frames`` frames, and its Es/N0 takes a reflected random walk
    of one bin (0.1 dB) per request over ``snr_bins`` bins."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per = rate_per_s / streams
    ts, ss, cs, fs, bs = [], [], [], [], []
    for s in range(streams):
        n = int(rng.poisson(per * duration_s * 1.2)) + 8
        t = np.cumsum(rng.exponential(1.0 / per, size=n)) if per > 0 else np.zeros(0)
        t = t[t < duration_s]
        steps = rng.choice((-1, 1), size=t.size)
        b = np.empty(t.size, dtype=np.int64)
        cur = int(rng.integers(0, snr_bins))
        for i, d in enumerate(steps):
            cur += int(d)
            if cur < 0:
                cur = 1
            elif cur >= snr_bins:
                cur = snr_bins - 2
            b[i] = cur
        ts.append(t)
        ss.append(np.full(t.size, s))
        cs.append(np.full(t.size, s % num_codes))
        fs.append(rng.integers(1, max_frames + 1, size=t.size))
        bs.append(b)
    t = np.concatenate(ts)
    order = np.argsort(t, kind="stable")
    return Schedule(t[order], np.concatenate(ss)[order], np.concatenate(cs)[order],
                    np.concatenate(fs)[order], np.concatenate(bs)[order])  # fmt: skip


def burst_schedule(requests: int, *, streams: int = 32, num_codes: int = 3, max_frames: int = 256,
                   snr_bins: int = 16, seed: int = 0) -> Schedule:
    """``requests`` requests all present at time 0 (saturation): streams
    round-robin, sizes uniform in ``1..max_frames``, Es/N0 bins uniform."""
    rng = np.random.Generator(np.random.PCG64(seed))
    stream = np.arange(requests) % streams
    return Schedule(np.zeros(requests), stream, stream % num_codes,
                    rng.integers(1, max_frames + 1, size=requests),
                    rng.integers(0, snr_bins, size=requests))  # fmt: skip


@dataclass
class LoadResult:
    """Outcome of :func:`drive`."""

    requests: int
    frames: int
    info_bits: int
    elapsed_s: float
    latency_ms: np.ndarray  # per frame
    rejected: int  # admissions refused at first try (budget full)
    launches: int
    extra: dict = field(default_factory=dict)

    @property
    def gbps(self) -> float:
        return self.info_bits / self.elapsed_s / 1e9

    def percentile(self, q: float) -> float:
        return float(np.percentile(self.latency_ms, q)) if self.latency_ms.size else float("nan")


def drive(server: Server, schedule: Schedule, fill: Callable[[Slot, int], None], *,
          clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] | None = None,
          idle_s: float = 0.0, on_done: Callable[[Slot, int], None] | None = None,
          admit_slice_s: float = 1e-3,
          before_step: Callable[[], None] | None = None) -> LoadResult:
    """Run ``schedule`` against ``server``: each request is admitted when it
    has arrived (or, when the budget is full, as soon as room frees up), its
    LLRs written by ``fill(slot, request index)``, submitted, and launched by
    the next :meth:`Server.step`; completions are polled every loop turn,
    handed to ``on_done(slot, request index)`` if given (work it enqueues on
    the server's stream reads the results before the slot can be reused) and
    released.  A loop turn admits for at most ``admit_slice_s`` seconds, so a
    long backlog is launched in pieces while the rest is still being admitted
    instead of leaving the device idle.  ``before_step()``, if given, runs
    after each turn's fills and before its :meth:`Server.step` (a fill that
    only queues its work, as :class:`~commkit_amd.mapping.demap.
    SegmentsDemapper.add` does, launches it there: the step's launch waits
    for the work enqueued on the caller's stream).  Python's cyclic garbage
    collector is paused while the loop runs (the loop makes no reference
    cycles; a full collection over a large heap stalls admission and polling
    for milliseconds) and restored after.  Returns per-frame latencies
    (completion seen minus arrival)."""
    gc_was_enabled = gc.isenabled()
    gc.disable()
    try:
        return _drive(server, schedule, fill, clock, sleep, idle_s, on_done, admit_slice_s,
                      before_step)
    finally:
        if gc_was_enabled:
            gc.enable()


def _drive(server: Server, schedule: Schedule, fill: Callable[[Slot, int], None],
           clock: Callable[[], float], sleep: Callable[[float], None] | None, idle_s: float,
           on_done: Callable[[Slot, int], None] | None, admit_slice_s: float,
           before_step: Callable[[], None] | None) -> LoadResult:
    n = len(schedule)
    t_arr = schedule.t
    t_done = np.full(n, np.nan)
    backlog: list[int] = []
    refused: set[int] = set()
    nxt = 0
    done = 0
    rejected = 0
    launches = 0
    start = clock()
    while done < n:
        now = clock() - start
        while nxt < n and t_arr[nxt] <= now:
            backlog.append(nxt)
            nxt += 1
        # First come, first served: admission stops at the first request that
        # does not fit, so a large request is not overtaken indefinitely.
        k = 0
        turn_end = clock() + admit_slice_s
        while k < len(backlog) and (k == 0 or clock() < turn_end):
            i = backlog[k]
            slot = server.admit(int(schedule.code[i]), int(schedule.frames[i]), i)
            if slot is None:
                if i not in refused:
                    refused.add(i)
                    rejected += 1
                break
            fill(slot, i)
            server.submit(i)
            k += 1
        backlog = backlog[k:]
        if before_step is not None:
            before_step()
        if server.step() > 0:
            launches += 1
        for tag in server.poll():
            t_done[tag] = clock() - start
            if on_done is not None:
                on_done(server.slot(tag), tag)
            server.release(tag)
            done += 1
        if sleep is not None and not backlog and nxt < n and done == nxt:
            wait = t_arr[nxt] - (clock() - start)
            if wait > idle_s:
                sleep(min(wait, 1e-3))
    elapsed = clock() - start
    lat = np.repeat((t_done - t_arr) * 1e3, schedule.frames)
    info = int(sum(server.codes[int(c)].k * int(f) for c, f in zip(schedule.code, schedule.frames)))
    return LoadResult(n, int(schedule.frames.sum()), info, elapsed, lat, rejected, launches)
