"""Batched decoding: frames of several QC codes in one persistent launch.

:class:`BatchDecoder` wraps ``ck_batch_*`` of ``libck_hip.so``.  The frames
of each code form a class (inputs ``(R_c, n_c)`` float32 LLRs on one GPU);
one persistent launch decodes every class, each workgroup refilling a frame
slot from its class's queue as soon as the slot's frame stops (syndrome
satisfied or the iteration limit), so frames that converge early do not hold
lanes while their neighbours iterate.  With ``graph=True`` the launch is
captured once in a HIP graph and replayed by :meth:`run`: the graph reads
and writes the same device buffers every time, so new frames are copied into
:attr:`inputs` in place between runs.

Every frame decodes to the same bits, APP LLRs, iteration count and valid
flag as the per-call :func:`commkit_amd.decode` of that frame alone.  This
has no counterpart in the reference, whose ``_decode`` launches one code's
frames in chunks (``commkit/coding/ldpc.py:565-581``).
"""

from __future__ import annotations

import ctypes
from typing import Any, Sequence

from . import _native
from .coding.ldpc import LDPCCode, MinSum, SumProduct, device_code, method_struct

__all__ = ["BatchDecoder", "BatchResult"]

MAX_CLASSES = 8


class CkBatchClass(ctypes.Structure):
    """``ck_batch_class`` of ``include/ck_tail.h``."""

    _fields_ = [
        ("code", ctypes.c_void_p),
        ("num_frames", ctypes.c_int64),
        ("llrs", ctypes.c_void_p),
        ("app", ctypes.c_void_p),
        ("bits", ctypes.c_void_p),
        ("iterations", ctypes.c_void_p),
        ("valid", ctypes.c_void_p),
        ("frames_per_group", ctypes.c_int32),
        ("lanes_per_row", ctypes.c_int32),
    ]


class CkBatchPlan(ctypes.Structure):
    """``ck_batch_plan`` of ``include/ck_tail.h``."""

    _fields_ = [
        ("workspace", ctypes.c_void_p),
        ("grid", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("sum_product", ctypes.c_int32),
        ("frames_per_group", ctypes.c_int32 * 8),
        ("lanes_per_row", ctypes.c_int32 * 8),
        ("kernel_variant", ctypes.c_int32),
    ]


class BatchResult:
    """Per-class outputs: ``bits`` (R, k) int8, ``llrs`` (R, n) float32,
    ``valid`` (R,) bool, ``iterations`` (R,) int32, device tensors."""

    def __init__(self, bits: Any, llrs: Any, valid: Any, iterations: Any) -> None:
        self.bits = bits
        self.llrs = llrs
        self.valid = valid
        self.iterations = iterations


class BatchDecoder:
    """One persistent launch decoding frames of up to 8 QC codes.

    Parameters
    ----------
    codes : sequence of LDPCCode
        One per class; each must have a quasi-cyclic view.  Classes run in
        the order given (put the largest frames first).
    num_frames : sequence of int
        Frames per class.
    method : MinSum or SumProduct
        Layered schedule.
    device : torch device, optional
        Default: the current GPU.
    graph : bool, default True
        Capture the launch in a HIP graph, replayed by :meth:`run`.
    tiles : sequence of (frames_per_group, lanes_per_row), optional
        Per class; 0 lets the library choose.
    """

    def __init__(self, codes: Sequence[LDPCCode], num_frames: Sequence[int],
                 method: MinSum | SumProduct, *, device: Any = None, graph: bool = True,
                 tiles: Sequence[tuple[int, int]] | None = None) -> None:
        import torch

        if not 1 <= len(codes) <= MAX_CLASSES or len(codes) != len(num_frames):
            raise ValueError(f"need 1 to {MAX_CLASSES} classes with a frame count each")
        if method.schedule != "layered":
            raise ValueError("batched decoding runs the layered schedule")
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.device = dev
        self.codes = list(codes)
        self.method = method
        self._lib = _native.lib()
        self._keep: list[Any] = []
        self.inputs: list[Any] = []
        self.outputs: list[BatchResult] = []
        arr = (CkBatchClass * len(codes))()
        with torch.cuda.device(dev):
            for i, (code, R) in enumerate(zip(codes, num_frames)):
                if code._qc_view is None:
                    raise ValueError(f"class {i}: {code!r} has no quasi-cyclic view")
                dc = device_code(code, dev)
                self._keep.append(dc)
                llrs = torch.zeros((int(R), code.n), dtype=torch.float32, device=dev)
                out = BatchResult(
                    torch.empty((int(R), code.k), dtype=torch.int8, device=dev),
                    torch.empty((int(R), code.n), dtype=torch.float32, device=dev),
                    torch.empty(int(R), dtype=torch.uint8, device=dev),
                    torch.empty(int(R), dtype=torch.int32, device=dev),
                )
                self.inputs.append(llrs)
                self.outputs.append(out)
                F, P = tiles[i] if tiles is not None else (0, 0)
                arr[i] = CkBatchClass(
                    ctypes.addressof(dc.struct), int(R), llrs.data_ptr(), out.llrs.data_ptr(),
                    out.bits.data_ptr(), out.iterations.data_ptr(), out.valid.data_ptr(), F, P,
                )  # fmt: skip
            self._classes = arr
            self._method = method_struct(codes[0], method)
            nbytes = self._lib.ck_batch_workspace_bytes(arr, len(codes), ctypes.byref(self._method))
            if nbytes < 0:
                plan = CkBatchPlan()
                _native.check(self._lib.ck_batch_prepare(arr, len(codes), ctypes.byref(self._method),
                                                         None, 0, ctypes.byref(plan)))  # fmt: skip
            self.workspace = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
            self.plan = CkBatchPlan()
            _native.check(self._lib.ck_batch_prepare(
                arr, len(codes), ctypes.byref(self._method), self.workspace.data_ptr(),
                int(nbytes), ctypes.byref(self.plan),
            ))  # fmt: skip
            self._graph = ctypes.c_void_p()
            if graph:
                _native.check(self._lib.ck_batch_graph_create(ctypes.byref(self.plan),
                                                              ctypes.byref(self._graph)))

    @property
    def frames_per_group(self) -> list[int]:
        return list(self.plan.frames_per_group)[: len(self.codes)]

    def run(self) -> list[BatchResult]:
        """Decode every frame of :attr:`inputs` (on torch's current stream)."""
        import torch

        with torch.cuda.device(self.device):
            stream = _native.stream_of(self.device)
            if self._graph:
                _native.check(self._lib.ck_batch_graph_launch(self._graph, stream))
            else:
                _native.check(self._lib.ck_batch_launch(ctypes.byref(self.plan), stream))
        return self.outputs

    def close(self) -> None:
        if getattr(self, "_graph", None):
            _native.check(self._lib.ck_batch_graph_destroy(self._graph))
            self._graph = ctypes.c_void_p()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass
