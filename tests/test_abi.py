"""The C-ABI library loads and exports every function include/*.h declares."""
import glob
import os
import re

from casl_p4est_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?\w+\s*\*?\s*(casl_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "casl_reinit" in names and "casl_stencil_create" in names and "casl_batch_reinit" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_bindings_cover_header():
    assert declared_functions() <= set(L.SIGNATURES), set(declared_functions()) - set(L.SIGNATURES)


def test_error_codes_on_bad_arguments():
    import ctypes
    h = ctypes.c_void_p()
    rc = L.lib().casl_forest_new(4, None, None, ctypes.byref(h))
    assert rc == L.CASL_ERR_INVALID_ARGUMENT
    assert b"dim" in L.lib().casl_last_error()
    rc = L.lib().casl_reinit(None, None, 2, 2, 1, 1.0, -1.0, 0, None)
    assert rc == L.CASL_ERR_INVALID_ARGUMENT


def test_device_cache_bookkeeping():
    """The cache's limit, held-bytes and trim calls; with nothing cached they
    touch no device, so they run here without a GPU."""
    import casl_p4est_amd as C
    assert C.device_cache_held(0) == 0
    C.device_cache_set_limit(1 << 30)
    assert C.device_cache_held(0) == 0
    C.device_cache_trim()
    C.device_cache_set_limit(0)
    assert L.lib().casl_device_cache_held(0, None) == L.CASL_ERR_INVALID_ARGUMENT
    import pytest
    with pytest.raises(ValueError):
        C.device_cache_set_limit(-1)
