"""FP64 oracle bindings -- TEST INFRASTRUCTURE ONLY.

ctypes front-end for ``oracle/build/libklt_oracle.so`` (the C restatement of
K-Litho's imaging path in ``klt_oracle.c``) and, when it was built, for
``oracle/_ref/libklref.so`` (K-Litho's own source/mask/file_io code).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU baseline import
this module; the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "build", "libklt_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libklref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference pieces when the tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        i, f, d = C.c_int, C.c_float, C.c_double
        sig = {
            "or_src_annular": (None, [_f32p, i, f, f]),
            "or_src_crossquad": (None, [_f32p, i, f, f]),
            "or_src_dipole": (None, [_f32p, i, f, f, i]),
            "or_src_point": (None, [_f32p, i, f, f]),
            "or_normalize": (None, [_f32p, i]),
            "or_mask_linespace": (None, [_f32p, i, i, i, i, i]),
            "or_mask_place": (None, [_f32p, i, i, _f32p, i, i, f]),
            "or_outer_sigma": (f, [_f32p, i]),
            "or_sizing": (None, [f, i, i, f, f, C.POINTER(i), C.POINTER(i)]),
            "or_count_source_points": (i, [_f32p, i, f]),
            "or_pupil_matrix": (i, [_f32p, i, f, f, f, f, i, i, i, i, f, _f64p, _f64p, _i32p]),
            "or_tcc": (None, [_f64p, _f32p, i, f, f, f, f, i, i, i, i, f]),
            "or_eigh_topk": (i, [_f64p, i, i, _f64p, _f64p]),
            "or_kernels_from_eig": (None, [_f64p, _f64p, i, i, _f32p, _f32p, _f32p]),
            "or_mask_band": (None, [_f32p, i, i, i, i, _f64p]),
            "or_socs_image": (None, [_f32p, i, i, _f32p, _f32p, _f32p, i, i, i, _f64p]),
            "or_socs_image_d": (None, [_f64p, i, i, _f32p, _f32p, _f32p, i, i, i, _f64p]),
            "or_hopkins_image": (None, [_f32p, i, i, _f64p, i, i, _f64p]),
            "or_socs_loss_grad": (d, [_f64p, _f32p, i, i, _f32p, _f32p, _f32p, i, i, i,
                                      d, d, C.c_void_p, C.c_void_p]),
            "or_abbe_loss_grad": (d, [_f64p, _f32p, i, i, _f32p, i, f, f, f, i, i, f, d, d,
                                      C.c_void_p, C.c_void_p]),
            "or_abbe_image": (None, [_f64p, i, i, _f32p, i, f, f, f, i, i, f, _f64p]),
            "or_resist_levels": (None, [_f32p, i, f, f, np.ctypeslib.ndpointer(np.uint8)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref():
    """K-Litho's own host code (None when oracle/_ref was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        R = C.CDLL(REF_SO)
        i, f, s = C.c_int, C.c_float, C.c_char_p
        R.ref_create_source.argtypes = [_f32p, i, s, f, f, f, f, i, f, f, s]
        R.ref_create_mask.argtypes = [_f32p, i, i, i, i, s, i, i, i, s, f]
        R.ref_write_kernel_info.argtypes = [s, i, i, i, _f32p]
        R.ref_read_kernel_info.argtypes = [s, C.POINTER(i), C.POINTER(i), _f32p, i]
        R.ref_read_kernel_info.restype = i
        R.ref_read_float_bin.argtypes = [s, _f32p, i]
        R.ref_read_float_bin.restype = i
        R.ref_write_float_bin.argtypes = [s, _f32p, i]
        for name in ("ref_png_gray", "ref_png_7colors", "ref_png_continuous"):
            getattr(R, name).argtypes = [s, i, i, _f32p, f, f]
        _ref = R
    return _ref


# --------------------------------------------------------------------------
# thin numpy front-ends
# --------------------------------------------------------------------------

def source(kind: str, n: int, inner=0.6, outer=0.9, radius=0.1, offset=0.5, onx=1,
           px=0.6, py=0.2, normalize=True) -> np.ndarray:
    """createSource (source.cpp:105-142) for the generated types."""
    m = np.zeros(n * n, np.float32)
    L = lib()
    if kind == "Annular":
        L.or_src_annular(m, n, inner, outer)
    elif kind == "Dipole":
        L.or_src_dipole(m, n, radius, offset, int(onx))
    elif kind == "CrossQuadrupole":
        L.or_src_crossquad(m, n, radius, offset)
    elif kind == "Point":
        L.or_src_point(m, n, px, py)
    if normalize:
        L.or_normalize(m, n * n)
    return m.reshape(n, n)


def outer_sigma(src: np.ndarray) -> float:
    s = np.ascontiguousarray(src, np.float32).ravel()
    return float(lib().or_outer_sigma(s, int(round(np.sqrt(s.size)))))


def sizing(NA: float, Lx: int, Ly: int, sigma: float, lam: float = 193.0):
    nx, ny = C.c_int(), C.c_int()
    lib().or_sizing(NA, Lx, Ly, sigma, lam, C.byref(nx), C.byref(ny))
    return nx.value, ny.value


def pupil_matrix(src, sigma, NA, defocus, Lx, Ly, Nx, Ny, lam=193.0, nmed=1.0):
    s = np.ascontiguousarray(src, np.float32).ravel()
    n = int(round(np.sqrt(s.size)))
    ns = lib().or_count_source_points(s, n, sigma)
    T = (2 * Nx + 1) * (2 * Ny + 1)
    pup = np.zeros((max(ns, 1), T, 2), np.float64)
    sv = np.zeros(max(ns, 1), np.float64)
    pq = np.zeros((max(ns, 1), 2), np.int32)
    lib().or_pupil_matrix(s, n, sigma, lam, NA, defocus, Lx, Ly, Nx, Ny, nmed, pup, sv, pq)
    return pup[:ns, :, 0] + 1j * pup[:ns, :, 1], sv[:ns], pq[:ns]


def tcc(src, sigma, NA, defocus, Lx, Ly, Nx, Ny, lam=193.0, nmed=1.0) -> np.ndarray:
    s = np.ascontiguousarray(src, np.float32).ravel()
    n = int(round(np.sqrt(s.size)))
    T = (2 * Nx + 1) * (2 * Ny + 1)
    out = np.zeros((T, T, 2), np.float64)
    lib().or_tcc(out, s, n, sigma, lam, NA, defocus, Lx, Ly, Nx, Ny, nmed)
    return out[..., 0] + 1j * out[..., 1]


def eigh_topk(a: np.ndarray, k: int):
    """Top-k eigenpairs, descending; vecs[i] is the eigenvector of w[i]."""
    n = a.shape[0]
    k = min(k, n)
    ai = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)
    w = np.zeros(k, np.float64)
    z = np.zeros((k, n, 2), np.float64)
    rc = lib().or_eigh_topk(ai, n, k, w, z)
    assert rc == 0
    return w, z[..., 0] + 1j * z[..., 1]


def kernels_from_eig(vecs: np.ndarray, w: np.ndarray):
    """calcKernels' reversal + float storage (klitho_tcc.cpp:343-362)."""
    k, T = vecs.shape
    v = np.ascontiguousarray(np.stack([vecs.real, vecs.imag], -1), np.float64)
    kr = np.zeros((k, T), np.float32)
    ki = np.zeros((k, T), np.float32)
    sc = np.zeros(k, np.float32)
    lib().or_kernels_from_eig(v, np.ascontiguousarray(w, np.float64), k, T, kr, ki, sc)
    return kr, ki, sc


def mask_band(mask: np.ndarray, Nx: int, Ny: int) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = m.shape
    out = np.zeros((2 * Ny + 1, 2 * Nx + 1, 2), np.float64)
    lib().or_mask_band(m, Lx, Ly, Nx, Ny, out)
    return out[..., 0] + 1j * out[..., 1]


def socs_image(mask, kr, ki, scales, Nx, Ny) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = m.shape
    kr = np.ascontiguousarray(kr, np.float32)
    ki = np.ascontiguousarray(ki, np.float32)
    sc = np.ascontiguousarray(scales, np.float32)
    img = np.zeros((Ly, Lx), np.float64)
    lib().or_socs_image(m, Lx, Ly, kr, ki, sc, int(sc.size), Nx, Ny, img)
    return img


def socs_image_d(mask, kr, ki, scales, Nx, Ny) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.float64)
    Ly, Lx = m.shape
    img = np.zeros((Ly, Lx), np.float64)
    lib().or_socs_image_d(m, Lx, Ly, np.ascontiguousarray(kr, np.float32),
                          np.ascontiguousarray(ki, np.float32),
                          np.ascontiguousarray(scales, np.float32), int(np.size(scales)),
                          Nx, Ny, img)
    return img


def hopkins_image(mask, tcc_mat, Nx, Ny) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.float32)
    Ly, Lx = m.shape
    t = np.ascontiguousarray(np.stack([tcc_mat.real, tcc_mat.imag], -1), np.float64)
    img = np.zeros((Ly, Lx), np.float64)
    lib().or_hopkins_image(m, Lx, Ly, t, Nx, Ny, img)
    return img


def socs_loss_grad(mask, target, kr, ki, scales, Nx, Ny, a=50.0, thr=0.25, grad=True):
    m = np.ascontiguousarray(mask, np.float64)
    Ly, Lx = m.shape
    t = np.ascontiguousarray(target, np.float32)
    img = np.zeros((Ly, Lx), np.float64)
    g = np.zeros((Ly, Lx), np.float64) if grad else None
    L = lib().or_socs_loss_grad(m, t, Lx, Ly, np.ascontiguousarray(kr, np.float32),
                                np.ascontiguousarray(ki, np.float32),
                                np.ascontiguousarray(scales, np.float32),
                                int(np.size(scales)), Nx, Ny, a, thr,
                                img.ctypes.data, g.ctypes.data if grad else None)
    return L, img, g


def abbe_loss_grad(mask, target, src, NA, defocus, Nx, Ny, a=50.0, thr=0.25, lam=193.0,
                   grad=True, nmed=1.0):
    m = np.ascontiguousarray(mask, np.float64)
    Ly, Lx = m.shape
    s = np.ascontiguousarray(src, np.float32)
    n = s.shape[0]
    t = np.ascontiguousarray(target, np.float32)
    img = np.zeros((Ly, Lx), np.float64)
    g = np.zeros((n, n), np.float64) if grad else None
    L = lib().or_abbe_loss_grad(m, t, Lx, Ly, s, n, lam, NA, defocus, Nx, Ny, nmed, a, thr,
                                img.ctypes.data, g.ctypes.data if grad else None)
    return L, img, g


def abbe_image(mask, src, NA, defocus, Nx, Ny, lam=193.0, nmed=1.0) -> np.ndarray:
    m = np.ascontiguousarray(mask, np.float64)
    Ly, Lx = m.shape
    s = np.ascontiguousarray(src, np.float32)
    img = np.zeros((Ly, Lx), np.float64)
    lib().or_abbe_image(m, Lx, Ly, s, s.shape[0], lam, NA, defocus, Nx, Ny, nmed, img)
    return img


def resist_loss(img: np.ndarray, target: np.ndarray, a=50.0, thr=0.25) -> float:
    s = 1.0 / (1.0 + np.exp(-a * (img.astype(np.float64) - thr)))
    return float(np.sum((s - target.astype(np.float64)) ** 2))


def resist_levels(img: np.ndarray, vmin: float = 0.0, vmax: float = 0.225) -> np.ndarray:
    """Class 0..6 of each pixel in the reference's 7-level view (file_io.cpp:25-80)."""
    d = np.ascontiguousarray(img, np.float32)
    out = np.empty(d.shape, np.uint8)
    lib().or_resist_levels(d.ravel(), d.size, vmin, vmax, out.reshape(-1))
    return out
