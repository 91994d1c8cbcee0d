"""The C-ABI library loads and exports every entry point include/icx.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re

import icx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = "".join(open(os.path.join(ROOT, "include", h)).read() for h in ("icx.h", "icx_profile.h"))
    return sorted(set(re.findall(r"^\s*(?:icx_status|const char \*)\s*(icx_\w+)\s*\(", txt, re.M)))


def test_header_and_python_lists_agree():
    assert declared() == sorted(icx.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(icx.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_context_create_reports_error_without_gpu():
    import torch
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = icx.lib().icx_ctx_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
