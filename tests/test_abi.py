"""The C-ABI library loads and exports every entry point include/klt.h declares."""
import ctypes as C
import os

import pytest

import klt


def test_library_exports_every_declared_symbol():
    L = klt.lib()
    names = klt.declared_symbols()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_errors_without_gpu():
    L = klt.lib()
    assert L.klt_version() >= 1
    n = L.klt_device_count()
    assert n >= 0
    if n == 0:
        h = C.c_void_p()
        rc = L.klt_ctx_create(0, C.byref(h))
        assert rc == -4  # KLT_ERR_NODEV
        assert b"device" in L.klt_last_error()


def test_null_arguments_are_rejected():
    L = klt.lib()
    assert L.klt_ctx_create(0, None) == -1
    assert L.klt_ctx_sync(None) == -1
    assert L.klt_socs_forward(None, None, None, 1, 8, 8, None, 0) == -1


def test_tools_are_built():
    for tool in ("klitho_tcc", "klitho_socs"):
        exe = os.path.join(klt.PKG, "bin", tool)
        if not os.path.exists(exe):
            pytest.skip("host tools not built yet")
        assert os.access(exe, os.X_OK)
