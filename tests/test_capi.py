"""Host-only checks of the C-ABI library: it loads, exports every entry point
include/mickks.h declares, and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import ROOT, product


def declared_symbols():
    with open(os.path.join(ROOT, "include", "mickks.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    M = product()
    lib = M.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(M.EXPORTED) == names


def test_context_rejects_bad_shapes_on_host():
    M = product()
    lib = M.lib()
    q = np.array([1152921504606584833], np.uint64)
    p = np.array([1152921504606748673], np.uint64)
    h = ctypes.c_void_p()
    rc = lib.mi_context_create(ctypes.byref(h), 1000, q.ctypes.data, 1, p.ctypes.data, 1, 1)
    assert rc == -1
    assert lib.mi_last_error().decode().startswith("mi_context_create: N must be")
    rc = lib.mi_context_create(ctypes.byref(h), 1024, q.ctypes.data, 1, p.ctypes.data, 1, 2)
    assert rc == -1 and "dnum" in lib.mi_last_error().decode()
    big = np.array([(1 << 63) + 1], np.uint64)
    rc = lib.mi_context_create(ctypes.byref(h), 1024, big.ctypes.data, 1, p.ctypes.data, 1, 1)
    assert rc == -1 and "2^62" in lib.mi_last_error().decode()
    assert lib.mi_ntt_forward(None, None, 1, 1, 1024, 0, None) == -1
    assert "null" in lib.mi_last_error().decode()
