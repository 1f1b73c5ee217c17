"""Signal storage: :func:`save_npz` and :func:`load_npz`.

The archive layout of ``commkit/io.py:1-43`` for what the receive tail
reads, so that archives pass between the two:

``samples``
    The samples (always present).
``reference_symbols``, ``reference_bits``
    The reference (omitted without one; each omitted when ``None``).
``reference_info_bits``, ``reference_code_block``
    The reference's source bits and the code's ``(n, k)`` (omitted when
    ``info_bits`` is ``bits``, as for an uncoded payload).
``reference_code.*``
    The array fields of the reference's ``code``, which is in
    ``__metadata__`` as its class name and fields.
``constellation_points``, ``constellation_bit_labels``, ``constellation_pmf``
    The constellation (omitted when ``None``; the pmf only when shaped).
``__metadata__``
    A 0-d unicode array holding a JSON string: the rates (``sampling_rate``
    / ``symbol_rate`` give samples per symbol), ``center_frequency``,
    ``domain``, the constellation family, the pulse (``None`` here) and the
    reference's code.

The archive holds only numeric and unicode arrays and is read with
``allow_pickle=False``: loading a file never executes code from it.  Codes
are rebuilt only from :class:`LDPCCode`, :class:`CCDM` and :class:`PAS`, by
calling their constructors with the stored fields; an LDPC code whose
parity checks are quasi-cyclic is rebuilt with its circulant structure, so a
loaded code takes the same decoder kernels as one built from its base
matrix.  Framed records (``__frame_metadata__``) are outside the receive
tail, which takes payload symbols only: they raise.
"""

from __future__ import annotations

import dataclasses
import json
import logging
from pathlib import Path
from typing import Any

import numpy as np

from .core import Reference, Signal
from .mapping.constellation import Constellation

__all__ = ["load_npz", "save_npz"]

logger = logging.getLogger("commkit_amd")


def _host(x: Any) -> Any:
    """A NumPy copy of a host array or device tensor (``None`` stays)."""
    if x is None:
        return None
    try:
        import torch

        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:  # pragma: no cover - torch is a dependency
        pass
    return np.asarray(x)


def save_npz(signal: Signal, path: str | Path, *, compressed: bool = True) -> None:
    """Save a :class:`Signal` to a NumPy archive (``.npz``), device arrays
    copied to the host first (``commkit/io.py:84-209``).

    ``path`` gets ``.npz`` appended when missing; ``compressed=False`` writes
    with ``savez`` (faster, larger).  Raises ValueError for a framed Signal
    and for a code that is not an LDPCCode, CCDM or PAS.
    """
    if not isinstance(signal, Signal):
        raise TypeError(f"save_npz(): expected a Signal, got {type(signal).__name__}.")
    if signal.frame is not None:
        raise ValueError("save_npz(): framed Signals are not stored; extract the payload first.")
    path = Path(path)
    if path.suffix != ".npz":
        path = path.with_name(path.name + ".npz")
    arrays: dict[str, Any] = {"samples": _host(signal.samples)}
    rate = 1.0 if signal.symbol_rate is None else float(signal.symbol_rate)
    meta: dict[str, Any] = {
        "sampling_rate": rate * signal.sps,
        "symbol_rate": rate,
        "center_frequency": 0.0,
        "domain": "time",
        "pulse": None,
        "reference_code": None,
    }
    ref = signal.reference
    if ref is not None:
        for key, value in (("reference_symbols", ref.symbols), ("reference_bits", ref.bits)):
            if value is not None:
                arrays[key] = _host(value)
        if ref.code is not None:
            info = ref.info_bits if ref.info_bits is not None else ref.bits
            if info is not None:
                arrays["reference_info_bits"] = _host(info)
            arrays["reference_code_block"] = np.array([ref.code.n, ref.code.k], dtype=np.int64)
            meta["reference_code"] = _code_to_form(ref.code, "reference_code", arrays)
    _put_constellation(signal.constellation, "constellation", arrays, meta)
    arrays["__metadata__"] = _json_array(meta)
    (np.savez_compressed if compressed else np.savez)(path, **arrays)
    logger.info("Saved Signal %s to %s.", arrays["samples"].shape, path)


def load_npz(path: str | Path, *, device: str = "cpu") -> Signal:
    """Load a :class:`Signal` written by :func:`save_npz` (or by commkit's,
    for a payload record), with NumPy arrays (``device="cpu"``) or tensors
    on the current GPU (``device="gpu"``) (``commkit/io.py:298-347``).

    Raises ValueError for an unknown device or code type, metadata that is
    not JSON, a pulse other than none, or a framed record.
    """
    if device not in ("cpu", "gpu"):
        raise ValueError(f"load_npz(): device must be 'cpu' or 'gpu', got {device!r}.")
    path = Path(path)
    if path.suffix != ".npz":
        path = path.with_name(path.name + ".npz")
    with np.load(path, allow_pickle=False) as data:
        sig = _read_signal(data, path)
    logger.info("Loaded Signal %s from %s, on %s.", sig.samples.shape, path, device)
    return sig if device == "cpu" else _to_gpu(sig)


def _read_signal(data: Any, path: Path) -> Signal:
    meta = _read_json(data, "__metadata__")
    if "__frame_metadata__" in data:
        raise ValueError(f"{path} holds a framed Signal; the receive tail takes payload "
                         "symbols only (extract the payload before saving).")  # fmt: skip
    if meta.get("pulse") is not None:
        raise ValueError(f"{path} holds pulse-shaped samples ({meta['pulse'].get('type')}); "
                         "the receive tail takes matched-filtered symbols at 1 sample "
                         "per symbol.")  # fmt: skip
    rs, sr = meta.get("sampling_rate"), meta.get("symbol_rate")
    sps = 1 if rs is None or sr is None else rs / sr
    if abs(sps - round(sps)) > 1e-9 or round(sps) < 1:
        raise ValueError(f"{path}: sampling_rate / symbol_rate = {sps} is not a whole number "
                         "of samples per symbol.")  # fmt: skip
    reference = None
    if "reference_symbols" in data or "reference_bits" in data:
        bits = data["reference_bits"] if "reference_bits" in data else None
        info = data["reference_info_bits"] if "reference_info_bits" in data else None
        reference = Reference(
            symbols=data["reference_symbols"] if "reference_symbols" in data else None,
            bits=bits,
            info_bits=info if info is not None else bits,
            code=_code_from_form(meta.get("reference_code"), data, path),
        )
    return Signal(
        samples=data["samples"],
        constellation=_get_constellation("constellation", data, meta),
        sps=int(round(sps)),
        reference=reference,
        symbol_rate=None if sr is None else float(sr),
    )


def _to_gpu(sig: Signal) -> Signal:
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("load_npz(device='gpu'): no GPU is visible to this process.")
    dev = torch.device("cuda", torch.cuda.current_device())

    def up(a: Any) -> Any:
        return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to(dev)

    ref = sig.reference
    if ref is not None:
        ref = Reference(symbols=up(ref.symbols), bits=up(ref.bits),
                        info_bits=up(ref.info_bits), code=ref.code)  # fmt: skip
    return dataclasses.replace(sig, samples=up(sig.samples), reference=ref)


def _put_constellation(c: Any, key: str, arrays: dict, meta: dict) -> None:
    meta[f"{key}_family"] = None if c is None else c.family
    if c is not None:
        arrays[f"{key}_points"] = c.points
        arrays[f"{key}_bit_labels"] = c.bit_labels
        if c.pmf is not None:
            arrays[f"{key}_pmf"] = c.pmf


def _get_constellation(key: str, data: Any, meta: dict) -> Any:
    if f"{key}_points" not in data:
        return None
    return Constellation(
        data[f"{key}_points"],
        bit_labels=data[f"{key}_bit_labels"],
        pmf=data[f"{key}_pmf"] if f"{key}_pmf" in data else None,
        family=meta.get(f"{key}_family"),
    )


def _code_types() -> dict[str, type]:
    from .coding import CCDM, PAS, LDPCCode

    return {cls.__name__: cls for cls in (LDPCCode, CCDM, PAS)}


# A code in the archive is a tree of JSON forms, the layout the reference's
# archives use, one form per kind of field (the key names the form):
#   None, bool, int, float, str   as themselves (NumPy scalars as Python's);
#   tuple                         {"tuple": [form, ...]};
#   ndarray                       {"array": name}, the array stored as `name`;
#   Constellation                 {"constellation": prefix, "family": family},
#                                 its arrays stored as `prefix_points`, ...;
#   LDPCCode, CCDM, PAS           {"type": class name, "fields": {name: form}},
#                                 over the constructor's fields.
# Array names extend the parent's name by the field path ("reference_code",
# then "reference_code.checks", "reference_code.matcher.composition", ...).


def _code_to_form(value: Any, name: str, arrays: dict) -> Any:
    """The JSON form of ``value`` (a code or a field of one); its arrays are
    added to ``arrays``."""
    if value is None or isinstance(value, (bool, int, float, str)):
        return value
    if isinstance(value, np.generic):
        return value.item()
    if isinstance(value, np.ndarray):
        arrays[name] = value
        return {"array": name}
    if isinstance(value, tuple):
        return {"tuple": [_code_to_form(item, f"{name}.{i}", arrays)
                          for i, item in enumerate(value)]}  # fmt: skip
    if isinstance(value, Constellation):
        meta: dict = {}
        _put_constellation(value, name, arrays, meta)
        return {"constellation": name, "family": meta[f"{name}_family"]}
    cls = type(value)
    if _code_types().get(cls.__name__) is not cls:
        raise ValueError(f"save_npz(): cannot store a {cls.__name__} in a code.")
    fields = {}
    for f in dataclasses.fields(value):
        if f.init:
            fields[f.name] = _code_to_form(getattr(value, f.name), f"{name}.{f.name}", arrays)
    return {"type": cls.__name__, "fields": fields}


def _code_from_form(form: Any, data: Any, path: Path) -> Any:
    """The value a JSON form of :func:`_code_to_form` stands for."""
    if not isinstance(form, dict):
        return form
    if "array" in form:
        return data[form["array"]]
    if "tuple" in form:
        return tuple(_code_from_form(item, data, path) for item in form["tuple"])
    if "constellation" in form:
        prefix = form["constellation"]
        return _get_constellation(prefix, data, {f"{prefix}_family": form["family"]})
    cls = _code_types().get(form.get("type"))
    if cls is None:
        # the reference's message (part of the load_npz contract)
        raise ValueError(f"Unknown code type {form.get('type')!r} in {path}.")
    kwargs = {k: _code_from_form(v, data, path) for k, v in form["fields"].items()}
    code = cls(**kwargs)
    return _quasi_cyclic(code) if cls.__name__ == "LDPCCode" else code


def _quasi_cyclic(code: Any) -> Any:
    """``code`` rebuilt from a base matrix when its parity checks are made of
    whole circulants (every block of ``Z x Z`` either empty or a shifted
    identity), for the largest such ``Z`` dividing both dimensions; the code
    itself otherwise.  The edges are identical either way."""
    from .coding import LDPCCode

    rows = code.checks.astype(np.int64)
    cols = code.variables.astype(np.int64)
    m, n = code.num_checks, code.num_variables
    g = int(np.gcd(m, n))
    for z in sorted((d for d in range(2, g + 1) if g % d == 0), reverse=True):
        block = (rows // z) * (n // z) + cols // z
        shift = (cols % z - rows % z) % z
        order = np.argsort(block, kind="stable")
        b, s = block[order], shift[order]
        starts = np.flatnonzero(np.r_[True, b[1:] != b[:-1]])
        counts = np.diff(np.r_[starts, b.size])
        if np.any(counts != z) or np.any(s != np.repeat(s[starts], counts)):
            continue
        base = np.full((m // z, n // z), -1, dtype=np.int64)
        base[b[starts] // (n // z), b[starts] % (n // z)] = s[starts]
        qc = LDPCCode.quasi_cyclic(base, lifting=z, punctured=code.punctured)
        if np.array_equal(qc.checks, code.checks) and np.array_equal(qc.variables,
                                                                      code.variables):
            return qc
    return code


def _json_array(obj: dict) -> np.ndarray:
    def scalar(value: Any) -> Any:
        if isinstance(value, np.generic):
            return value.item()
        raise TypeError(f"Cannot store {type(value).__name__} in npz metadata.")

    return np.array(json.dumps(obj, default=scalar))


def _read_json(data: Any, key: str) -> dict:
    try:
        return json.loads(str(data[key]))
    except (KeyError, ValueError) as exc:
        raise ValueError(f"{key!r} in this archive is missing or not JSON; archives with "
                         "pickled metadata are not loaded, since unpickling can execute "
                         "code.") from exc
