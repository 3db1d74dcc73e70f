"""The C-ABI library loads and exports every entry point its header
declares (CPU: no device calls)."""
import ctypes
import os
import re

import quids_amd
from quids_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "quids_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qd_\w+)\(", txt, re.M)))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_python_binding_covers_header():
    assert set(declared()) == set(_abi.SIGNATURES)


def test_version_and_status_strings():
    lib = _abi.lib()
    assert lib.qd_abi_version() == 2
    assert lib.qd_status_string(0) == b"ok"
    assert lib.qd_status_string(5) == b"unsupported"


def test_errors_are_reported():
    lib = _abi.lib()
    assert lib.qd_ctx_create(None, 0) == _abi.QD_ERR_INVALID
    assert b"null" in lib.qd_last_error()
    r = _abi.QdRule(kind=99)
    nc, ms = ctypes.c_uint32(), ctypes.c_uint32()
    assert lib.qd_rule_num_child(ctypes.byref(r), b"abcd", 4, ctypes.byref(nc), ctypes.byref(ms)) == \
        _abi.QD_ERR_UNSUPPORTED


def test_header_rule_kinds_match_binding():
    txt = open(os.path.join(ROOT, "include", "quids_hip.h")).read()
    for name in ("QD_RULE_HADAMARD", "QD_RULE_QCGD_SPLIT_MERGE", "QD_MOD_QCGD_STEP", "QD_MOD_ZGATE"):
        val = int(re.search(name + r"\s*=\s*(\d+)", txt).group(1))
        assert getattr(_abi, name) == val
