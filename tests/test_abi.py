"""The C-ABI library loads and exports every symbol include/marqov_amd.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "marqov_amd.h")).read()
    return sorted(set(re.findall(r"^MQ_API [^(]*?\b(mq_\w+)\(", txt, flags=re.M)))


def test_library_exports_all_declared_symbols():
    from marqov_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    from marqov_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()
    _lib.load()


def test_ctx_create_without_device_fails_cleanly():
    import pytest
    from marqov_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.mq_ctx_create(0, ctypes.byref(h))
    if rc == 0:  # a device is present: nothing to check here
        lib.mq_ctx_destroy(h)
        pytest.skip("a device is present")
    assert not h.value
