"""The reference's CPU path (FFTW3 + zheevr_) for bench.py's CPU baseline --
BASELINE AND TEST INFRASTRUCTURE ONLY.

ctypes front-end of ``oracle/build/libklt_cpuref.so`` (``cpu_ref_path.cpp``),
which takes the FFTW3 interface and zheevr_ at run time from MKL's
``libmkl_rt.so`` (``KLT_CPUREF_LIB`` names another library exporting them).
``available()`` is False, with the reason in ``reason()``, when no such
library loads; nothing is substituted then.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "build", "libklt_cpuref.so")
CANDIDATES = ["/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"]

_lib = None
_status = None  # (ok, text): library version or the reason it is unavailable

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _load():
    global _lib, _status
    if _status is not None:
        return
    if not os.path.exists(SO):
        _status = (False, f"unavailable: {SO} not built")
        return
    L = C.CDLL(SO)
    L.cr_load.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    i = C.c_int
    L.cr_socs_images.argtypes = [_f32p, i, i, i, _f32p, _f32p, _f32p, i, i, i, i, _f32p]
    L.cr_zheevr_topk.argtypes = [_f64p, i, i, i, _f64p, _f64p]
    # the library's OpenMP threading layer: GNU, like the rest of the process
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    paths = [os.environ["KLT_CPUREF_LIB"]] if os.environ.get("KLT_CPUREF_LIB") else CANDIDATES
    msg = C.create_string_buffer(512)
    why = "unavailable: no FFTW3 / LAPACK library found"
    for p in paths:
        if not os.path.exists(p):
            continue
        if L.cr_load(p.encode(), msg, len(msg)) == 0:
            _lib = L
            _status = (True, msg.value.decode(errors="replace").strip())
            return
        why = msg.value.decode(errors="replace")
    _status = (False, why)


def available() -> bool:
    _load()
    return _status[0]


def reason() -> str:
    """The library's version string, or why it is unavailable."""
    _load()
    return _status[1]


def socs_images(masks, kr, ki, scales, Nx, Ny, threads=1) -> np.ndarray:
    """FT_r2c + calcSOCS + FI of masks [B, Ly, Lx] with `threads` workers."""
    _load()
    if not _status[0]:
        raise RuntimeError(_status[1])
    m = np.ascontiguousarray(masks, np.float32)
    if m.ndim == 2:
        m = m[None]
    B, Ly, Lx = m.shape
    out = np.empty_like(m)
    sc = np.ascontiguousarray(scales, np.float32)
    rc = _lib.cr_socs_images(m, B, Lx, Ly, np.ascontiguousarray(kr, np.float32),
                             np.ascontiguousarray(ki, np.float32), sc, int(sc.size), Nx, Ny,
                             int(threads), out)
    if rc != 0:
        raise ValueError("ERROR in FI()" if rc == -2 else f"cr_socs_images: {rc}")
    return out


def zheevr_topk(a: np.ndarray, k: int, threads=1):
    """calcKernels' zheevr_ call: (w [k] descending, vecs [k, n] complex)."""
    _load()
    if not _status[0]:
        raise RuntimeError(_status[1])
    n = a.shape[0]
    ai = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)
    w = np.zeros(k, np.float64)
    z = np.zeros((k, n, 2), np.float64)
    rc = _lib.cr_zheevr_topk(ai, n, k, int(threads), w, z)
    if rc != 0:
        raise RuntimeError(f"zheevr_ info {rc}")
    return w, z[..., 0] + 1j * z[..., 1]
