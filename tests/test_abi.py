"""The native library loads and exports every function include/hyquas_amd.h
declares (no GPU calls), and the ctypes layer declares them all."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "hyquas_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hq_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported(hq):
    names = _declared()
    assert len(names) > 40
    lib = ctypes.CDLL(hq.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header(hq):
    hq.lib()
    assert set(_declared()) <= set(hq.EXPORTED)


def test_struct_layout(hq):
    assert ctypes.sizeof(hq.HQGate) == 80
    assert hq.HQGate.mat.offset == 16


def test_noalloc_entry_point_rejects_null_state(hq):
    # argument checks come before any HIP call: no GPU needed
    z = (ctypes.c_double * 4)()
    rc = hq.lib().hq_expect_z_noalloc(None, z, None)
    assert rc == hq.HQ_ERR_ARG
    assert b"null state" in hq.lib().hq_last_error()
