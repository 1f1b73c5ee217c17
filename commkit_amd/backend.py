"""Where the data lives, and the per-object device-table cache.

The receive tail runs on one GPU through hand-written HIP kernels.  Input
arrives in one of two forms:

- a device array: a PyTorch-ROCm tensor on a GPU, or any object exporting
  ``__dlpack__`` for a GPU (it is wrapped by ``torch.from_dlpack`` without a
  copy).  The kernels run on that tensor's device, on torch's current stream
  of that device, and the outputs are tensors on the same device
  ("the device follows the data", ``commkit/backend.py:290-333``);
- a host array: NumPy (or a CPU tensor, or a list).  It is copied to the
  current GPU, computed by the same kernels and copied back, so the outputs
  are host arrays of the input's kind.  There is no CPU compute path.

Tables built from value objects (constellation levels, code schedules) are
uploaded once per object and device and cached weakly on the object, as the
reference's ``_device_tables`` (``commkit/backend.py:256-287``).
"""

from __future__ import annotations

import weakref
from dataclasses import dataclass
from typing import Any, Callable

import numpy as np

__all__ = ["Placed", "dispatch", "device_tables", "to_host_like"]


def _torch():
    import torch

    return torch


@dataclass
class Placed:
    """An input moved to the compute device.

    ``tensor`` is a torch tensor on a GPU; ``kind`` says how to hand results
    back: ``"device"`` (keep tensors), ``"numpy"`` or ``"torch_cpu"``.
    """

    tensor: Any
    kind: str

    @property
    def device(self) -> Any:
        return self.tensor.device


def is_device_array(x: Any) -> bool:
    torch = _torch()
    return isinstance(x, torch.Tensor) and x.device.type == "cuda"


def dispatch(x: Any, name: str = "array") -> Placed:
    """Put ``x`` on a GPU as a torch tensor and remember where it came from."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if x.device.type == "cuda":
            return Placed(x, "device")
        return Placed(_upload(x), "torch_cpu")
    if hasattr(x, "__dlpack__") and not isinstance(x, np.ndarray):
        dev = x.__dlpack_device__() if hasattr(x, "__dlpack_device__") else (1, 0)
        t = torch.from_dlpack(x)
        if t.device.type == "cuda":
            return Placed(t, "device")
        return Placed(_upload(t), "numpy" if int(dev[0]) == 1 else "torch_cpu")
    arr = np.asarray(x)
    if arr.dtype == object:
        raise TypeError(f"{name}: unsupported array type {type(x).__name__}.")
    return Placed(_upload(torch.from_numpy(np.ascontiguousarray(arr))), "numpy")


def _upload(t: Any) -> Any:
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "commkit_amd runs the receive tail on a GPU through its HIP "
            "library; no GPU is visible to this process."
        )
    return t.to(device=torch.device("cuda", torch.cuda.current_device()))


def to_host_like(t: Any, kind: str) -> Any:
    """Hand a device result back in the caller's form."""
    if t is None or kind == "device":
        return t
    host = t.cpu()
    if kind == "torch_cpu":
        return host
    return host.numpy()


_TABLES: "weakref.WeakKeyDictionary[Any, dict]" = weakref.WeakKeyDictionary()


def device_tables(
    owner: Any, name: str, device: Any, build: Callable[[], tuple[np.ndarray, ...]]
) -> tuple[Any, ...]:
    """Host tables from ``build()`` as device tensors, cached per owner and device.

    ``owner`` must be weak-referenceable and hashable (the value objects
    are); ``None`` disables caching.  Arrays are uploaded contiguous, with
    their dtype.
    """
    torch = _torch()
    key = (name, str(device))
    if owner is not None:
        cache = _TABLES.setdefault(owner, {})
        hit = cache.get(key)
        if hit is not None:
            return hit
    # np.array copies: read-only host tables (frozen value objects) would
    # otherwise be wrapped as non-writable tensors before the upload.
    tables = tuple(torch.from_numpy(np.array(a, order="C")).to(device) for a in build())
    if owner is not None:
        cache[key] = tables
    return tables
