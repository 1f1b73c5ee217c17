"""The C ABI of libck_hip.so on the CPU: it loads, exports every function
include/ck_tail.h declares with the argument count the ctypes binding uses,
and rejects bad arguments with a status and a message before any launch."""

import ctypes
import os
import re

import pytest

from commkit_amd import _native
from commkit_amd.coding.ldpc import CkLdpcCode, CkLdpcMethod
from commkit_amd.mapping.demap import CkAxis, CkLlrSegment

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ck_tail.h")


def _declared() -> dict[str, int]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int32_t|int64_t|const char\*)\s+(ck_\w+)\s*\(([^)]*)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_header_and_binding_agree():
    declared = _declared()
    assert set(declared) == set(_native.SYMBOLS), (set(declared) ^ set(_native.SYMBOLS))
    for name, nargs in declared.items():
        assert len(_native.SYMBOLS[name][1]) == nargs, name


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.ck_version() >= 10000


def test_struct_layouts_match_header():
    # ck_axis: 6 pointers + 2 int32; ck_ldpc_method: 10 four-byte fields.
    assert ctypes.sizeof(CkAxis) == 6 * 8 + 8
    assert ctypes.sizeof(CkLdpcMethod) == 40
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct ck_ldpc_code \{(.*?)\} ck_ldpc_code;", text, re.S).group(1)
    fields = re.findall(r"(\w+)\s*;", body)
    assert [f[0] for f in CkLdpcCode._fields_] == fields
    body = re.search(r"typedef struct ck_llr_segment \{(.*?)\} ck_llr_segment;", text,
                     re.S).group(1)  # fmt: skip
    assert [f[0] for f in CkLlrSegment._fields_] == re.findall(r"(\w+)\s*;", body)
    assert ctypes.sizeof(CkLlrSegment) == 9 * 8


def test_argument_errors_return_status_and_message():
    lib = _native.lib()
    st = lib.ck_llr_separable_maxlog(None, 10, None, None, ctypes.c_double(1.0), None, None)
    assert st == 1 and b"axis is NULL" in lib.ck_last_error()
    ax = CkAxis(1, 1, 1, 1, 1, 1, 8, 3)
    st = lib.ck_llr_separable(None, 4, ctypes.byref(ax), ctypes.byref(ax), ctypes.c_double(1.0), 2,
                              None, None)  # fmt: skip
    assert st == 1 and b"exact must be 0 or 1" in lib.ck_last_error()
    st = lib.ck_llr_points(None, 4, None, None, None, 16, 3, ctypes.c_double(1.0), 0, None, None)
    assert st == 1 and b"M = 16 != 2^k" in lib.ck_last_error()
    st = lib.ck_ldpc_decode(ctypes.byref(CkLdpcCode()), None, None, None, None, None, None, 1,
                            None, None)  # fmt: skip
    assert st == 1 and b"num_checks" in lib.ck_last_error()
    assert lib.ck_ldpc_workspace_bytes(ctypes.byref(CkLdpcCode()), None, 1) == -1
    # segments are all checked before anything is launched
    segs = (CkLlrSegment * 2)(CkLlrSegment(1, 1, 10, 1.0, None, 1, None, 1, 0),
                              CkLlrSegment(1, 1, 10, 1.0, None, 1, 1, 7, 20))
    st = lib.ck_llr_separable_segments(segs, 2, ctypes.byref(ax), ctypes.byref(ax), 0, 0, None)
    assert st == 1 and b"segment 1: limit 20 is not a whole number of blocks of 7" in \
        lib.ck_last_error()
    segs[1] = CkLlrSegment(1, 1, 10, 0.0, None, 1, None, 1, 0)
    st = lib.ck_llr_separable_segments(segs, 2, ctypes.byref(ax), ctypes.byref(ax), 0, 0, None)
    assert st == 1 and b"segment 1: inv_s2 must be positive" in lib.ck_last_error()


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "_LIB", None)
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "libck_hip.so"))
    with pytest.raises(RuntimeError, match="not built"):
        _native.lib()
