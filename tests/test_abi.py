"""The C ABI: libmgfe.so exists, loads, and exports every entry point
include/mgfe.h declares (no compute call: this runs without a GPU)."""
import ctypes
import os
import re

from morganapublic_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mgfe.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    L = _lib.lib()
    for name in header_symbols():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_no_context_without_gpu_is_an_error_not_a_fallback():
    # on a machine without a GPU the context must fail loudly
    import torch
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    rc = _lib.lib().mg_ctx_create(0, ctypes.byref(h))
    assert rc != _lib.MG_OK and not h.value
