"""ctypes binding of libklt.so (the C ABI declared in include/klt.h).

Array arguments may be numpy arrays (host memory: passed with KLT_HOST_PTRS)
or torch tensors on the GPU (device memory: passed as raw pointers).  Every
call raises KltError with the library's message on a non-zero status.  The
library must have been built (k-litho_amd/build.py); there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libklt.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "klt.h")

KLT_OK = 0
KLT_HOST_PTRS = 1
KLT_EIG_PSD = 2
KLT_EIG_WARM = 4

ERRORS = {-1: "KLT_ERR_ARG", -2: "KLT_ERR_SIZE", -3: "KLT_ERR_HIP", -4: "KLT_ERR_NODEV",
          -5: "KLT_ERR_IO", -6: "KLT_ERR_NOMEM", -7: "KLT_ERR_EIG"}


class KltError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg


_lib = None


def lib() -> C.CDLL:
    """Load libklt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run k-litho_amd/build.py")
        try:
            # torch ships its own libamdhip64.so.7; loading it first makes
            # libklt bind to the same HIP runtime, so one process never holds
            # two runtimes (torch then reports no GPU).
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, i, u, fp = C.c_void_p, C.c_int, C.c_uint, C.c_void_p
        L.klt_last_error.restype = C.c_char_p
        L.klt_version.restype = i
        L.klt_device_count.restype = i
        L.klt_ctx_create.argtypes = [i, C.POINTER(vp)]
        L.klt_ctx_destroy.argtypes = [vp]
        L.klt_ctx_destroy.restype = None
        L.klt_ctx_set_stream.argtypes = [vp, vp]
        L.klt_ctx_stream.argtypes = [vp]
        L.klt_ctx_stream.restype = vp
        L.klt_ctx_sync.argtypes = [vp]
        L.klt_ctx_wait_stream.argtypes = [vp, vp]
        L.klt_ctx_join_stream.argtypes = [vp, vp]
        L.klt_ctx_set_profiling.argtypes = [vp, i]
        L.klt_ctx_profile_report.argtypes = [vp, C.c_char_p, i]
        L.klt_ctx_set_autotune.argtypes = [vp, i]
        L.klt_ctx_set_splits.argtypes = [vp, i, i, i]
        L.klt_ctx_set_band_rows.argtypes = [vp, i]
        L.klt_ctx_tune_report.argtypes = [vp, C.c_char_p, i]
        f = C.c_float
        cs = C.c_char_p
        L.klt_source_create.argtypes = [fp, i, cs, f, f, f, f, i, f, f, cs]
        L.klt_mask_create.argtypes = [fp, i, i, i, i, cs, i, i, i, cs, f]
        L.klt_outer_sigma.argtypes = [fp, i]
        L.klt_outer_sigma.restype = f
        L.klt_tcc_sizing.argtypes = [f, i, i, f, f, C.POINTER(i), C.POINTER(i)]
        L.klt_read_floats.argtypes = [cs, fp, i]
        L.klt_write_floats.argtypes = [cs, fp, i]
        L.klt_write_kernel_info.argtypes = [cs, i, i, i, fp]
        L.klt_read_kernel_info.argtypes = [cs, C.POINTER(i), C.POINTER(i), fp, i]
        L.klt_tcc_build.argtypes = [vp, fp, i, f, f, f, f, f, i, i, i, i, i, fp, u]
        L.klt_eig_topk.argtypes = [vp, fp, i, i, fp, fp, u]
        L.klt_eig_topk_batch.argtypes = [vp, fp, i, i, i, fp, fp, u]
        L.klt_calc_kernels.argtypes = [vp, fp, i, i, fp, fp, fp, u]
        L.klt_hopkins_image.argtypes = [vp, fp, i, i, fp, i, i, fp, u]
        L.klt_socs_backward.argtypes = [vp, vp, fp, fp, i, i, i, C.c_float, C.c_float, fp, fp,
                                        fp, u]
        L.klt_socs_vjp.argtypes = [vp, vp, fp, fp, i, i, i, fp, u]
        L.klt_abbe_source_grad.argtypes = [vp, fp, fp, i, i, i, fp, i, f, f, f, f, i, i, f, f,
                                           fp, fp, fp, u]
        L.klt_kernels_create.argtypes = [vp, fp, fp, fp, i, i, i, u, C.POINTER(vp)]
        L.klt_kernels_destroy.argtypes = [vp]
        L.klt_kernels_destroy.restype = None
        L.klt_socs_image.argtypes = [vp, vp, fp, i, i, fp, u]
        L.klt_socs_forward.argtypes = [vp, vp, fp, i, i, i, fp, u]
        L.klt_socs_forward_noalloc.argtypes = [vp, vp, fp, i, i, i, fp]
        L.klt_socs_forward_batch.argtypes = [vp, i, vp, vp, vp, i, vp, vp, u]
        L.klt_socs_batch_latency.argtypes = [vp, fp, i]
        for name in ("klt_write_png_gray", "klt_write_png_7colors", "klt_write_png_continuous"):
            getattr(L, name).argtypes = [cs, i, i, fp, f, f]
        L.klt_resist_levels.argtypes = [vp, fp, i, i, i, f, f, fp, u]
        L.klt_smo_run.argtypes = [vp, vp, vp, i, i, i, fp, i, f, f, f, f, i, f, f, f, f, i, fp,
                                  fp]
        _lib = L
    return _lib


def check(rc: int):
    if rc != KLT_OK:
        raise KltError(rc, lib().klt_last_error().decode(errors="replace"))


def declared_symbols() -> list[str]:
    """Names of the functions include/klt.h declares."""
    import re
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(klt_[a-z0-9_]+)\s*\(", txt)))


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a, dtype):
    """(pointer, is_host) for a numpy array or a torch GPU tensor."""
    if _is_torch(a):
        import torch
        want = {np.float32: torch.float32, np.float64: torch.float64,
                np.int32: torch.int32}[dtype]
        if not a.is_cuda or a.dtype != want or not a.is_contiguous():
            raise TypeError(f"expected a contiguous {want} GPU tensor")
        return C.c_void_p(a.data_ptr()), False
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous:
        raise TypeError(f"expected a C-contiguous numpy {np.dtype(dtype)} array")
    return a.ctypes.data_as(C.c_void_p), True


class _TorchOrder:
    """Stream order around a call that takes torch tensors.  The library works
    on its own stream; this makes the call start after the work queued so far
    on torch's current stream (the kernels that produced its inputs) and
    torch's later work start after the call (the consumers of its outputs)."""

    def __init__(self, handle, *arrays):
        self.h, self.s = handle, None
        if any(_is_torch(a) for a in arrays if a is not None):
            import torch
            self.s = C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def __enter__(self):
        if self.s is not None:
            check(lib().klt_ctx_wait_stream(self.h, self.s))
        return self

    def __exit__(self, et, ev, tb):
        if self.s is not None and et is None:
            check(lib().klt_ctx_join_stream(self.h, self.s))
        return False


def _flags(*hosts) -> int:
    if all(hosts):
        return KLT_HOST_PTRS
    if any(hosts):
        raise TypeError("mix of host and device arrays in one call")
    return 0


def _cstr(s) -> bytes:
    return (s or "").encode() if isinstance(s, str) or s is None else s


def source_create(src_size: int, src_type: str, inner=0.6, outer=0.9, radius=0.1, offset=0.5,
                  on_x_axis=1, pt_x=0.6, pt_y=0.2, source_file="") -> np.ndarray:
    """createSource (source.cpp:105-142) through the library."""
    out = np.zeros((src_size, src_size), np.float32)
    check(lib().klt_source_create(out.ctypes.data_as(C.c_void_p), src_size, _cstr(src_type),
                                  inner, outer, radius, offset, int(on_x_axis), pt_x, pt_y,
                                  _cstr(source_file)))
    return out


def mask_create(Lx: int, Ly: int, msx: int, msy: int, mask_type: str, line_width=128,
                space_width=128, is_horizontal=0, mask_file="", dose=1.0) -> np.ndarray:
    """createMask (mask.cpp:31-60) through the library."""
    out = np.zeros((Ly, Lx), np.float32)
    check(lib().klt_mask_create(out.ctypes.data_as(C.c_void_p), Lx, Ly, msx, msy,
                                _cstr(mask_type), line_width, space_width, int(is_horizontal),
                                _cstr(mask_file), dose))
    return out


def outer_sigma(src: np.ndarray) -> float:
    s = np.ascontiguousarray(src, np.float32)
    return float(lib().klt_outer_sigma(s.ctypes.data_as(C.c_void_p), s.shape[0]))


def tcc_sizing(NA: float, Lx: int, Ly: int, sigma: float, lam: float = 193.0):
    nx, ny = C.c_int(), C.c_int()
    check(lib().klt_tcc_sizing(NA, Lx, Ly, sigma, lam, C.byref(nx), C.byref(ny)))
    return nx.value, ny.value


TARGET_INTENSITY = 0.225  # klitho_tcc.cpp:442


def write_png(kind: str, path: str, data: np.ndarray, vmin: float, vmax: float):
    """The reference's PNG views (file_io.cpp:7-135): kind 'gray', '7colors' or
    'continuous'; data [sy, sx] float32."""
    d = np.ascontiguousarray(data, np.float32)
    sy, sx = d.shape
    fn = getattr(lib(), f"klt_write_png_{kind}")
    check(fn(_cstr(path), sx, sy, d.ctypes.data_as(C.c_void_p), vmin, vmax))


def read_floats(path: str, size: int) -> np.ndarray:
    out = np.zeros(size, np.float32)
    check(lib().klt_read_floats(_cstr(path), out.ctypes.data_as(C.c_void_p), size))
    return out


def write_floats(path: str, data: np.ndarray):
    d = np.ascontiguousarray(data, np.float32).ravel()
    check(lib().klt_write_floats(_cstr(path), d.ctypes.data_as(C.c_void_p), d.size))


def write_kernel_info(path: str, size_x: int, size_y: int, scales: np.ndarray):
    s = np.ascontiguousarray(scales, np.float32)
    check(lib().klt_write_kernel_info(_cstr(path), size_x, size_y, s.size,
                                      s.ctypes.data_as(C.c_void_p)))


def read_kernel_info(path: str, cap: int = 4096):
    sx, sy = C.c_int(), C.c_int()
    sc = np.zeros(cap, np.float32)
    nk = lib().klt_read_kernel_info(_cstr(path), C.byref(sx), C.byref(sy),
                                    sc.ctypes.data_as(C.c_void_p), cap)
    if nk < 0:
        check(nk)
    return sx.value, sy.value, sc[:nk].copy()


class Context:
    """One libklt context (device, stream, plans, scratch)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().klt_ctx_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().klt_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream_ptr: int | None):
        check(lib().klt_ctx_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    def stream(self) -> int:
        return lib().klt_ctx_stream(self._h) or 0

    def sync(self):
        check(lib().klt_ctx_sync(self._h))

    def set_profiling(self, on: bool):
        check(lib().klt_ctx_set_profiling(self._h, 1 if on else 0))

    def set_autotune(self, on: bool):
        check(lib().klt_ctx_set_autotune(self._h, 1 if on else 0))

    def set_splits(self, ks: int = 0, rs: int = 0, kg: int = 0):
        """Pins the band column splits, synthesis row splits and kernel groups
        of later batched calls (0: automatic)."""
        check(lib().klt_ctx_set_splits(self._h, ks, rs, kg))

    def set_band_rows(self, rows: int = 0):
        """Pins the rows per block of the mask band pass of later batched calls
        (64 or 128; 0: automatic)."""
        check(lib().klt_ctx_set_band_rows(self._h, rows))

    def tune_report(self) -> list:
        """[(kind, B, Lx, Ly, Nx, Ny, nk, column_splits, row_splits, kernel_groups,
        band_rows, ms)] of the tuned shapes."""
        buf = C.create_string_buffer(1 << 16)
        n = lib().klt_ctx_tune_report(self._h, buf, len(buf))
        if n < 0:
            check(n)
        out = []
        for line in buf.value.decode().splitlines():
            f = line.split()
            out.append((f[0], *map(int, f[1:11]), float(f[11])))
        return out

    def profile_report(self) -> dict:
        """{kernel name: (calls, total ms)} since the last report."""
        buf = C.create_string_buffer(1 << 16)
        n = lib().klt_ctx_profile_report(self._h, buf, len(buf))
        if n < 0:
            check(n)
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, ms = line.split()
            out[name] = (int(calls), float(ms))
        return out

    def tcc_build(self, src, sigma: float, NA: float, defocus: float, Lx: int, Ly: int,
                  Nx: int, Ny: int, precision: int = 0, nmed: float = 1.0, lam: float = 193.0,
                  out=None):
        """calcTCC on the GPU; precision 0 = FP64 Gram, 1 = split-FP16 MFMA."""
        T = (2 * Nx + 1) * (2 * Ny + 1)
        if out is None:
            if _is_torch(src):
                import torch
                out = torch.empty((T, T, 2), dtype=torch.float64, device=src.device)
            else:
                out = np.empty((T, T, 2), np.float64)
        ps, hs = _ptr(src, np.float32)
        po, ho = _ptr(out, np.float64)
        n = int(src.shape[0])
        with _TorchOrder(self._h, src, out):
            check(lib().klt_tcc_build(self._h, ps, n, sigma, lam, NA, defocus, nmed, Lx, Ly, Nx,
                                      Ny, precision, po, _flags(hs, ho)))
        return out

    def eig_topk(self, a, k: int, psd: bool = False, warm: bool = False):
        """Top-k eigenpairs of a Hermitian matrix a [n, n, 2] (re, im), FP64 on the GPU;
        psd=True: a is positive semidefinite (KLT_EIG_PSD); warm: start from the
        previous solve's block (KLT_EIG_WARM).

        Returns (w [k], vecs [k, n, 2]) on the same side (host/device) as a."""
        n = int(a.shape[0])
        k = min(k, n)
        if _is_torch(a):
            import torch
            w = torch.empty(k, dtype=torch.float64, device=a.device)
            v = torch.empty((k, n, 2), dtype=torch.float64, device=a.device)
        else:
            w = np.empty(k, np.float64)
            v = np.empty((k, n, 2), np.float64)
        pa, ha = _ptr(a, np.float64)
        pw, hw = _ptr(w, np.float64)
        pv, hv = _ptr(v, np.float64)
        with _TorchOrder(self._h, a):
            check(lib().klt_eig_topk(self._h, pa, n, k, pw, pv,
                                     _flags(ha, hw, hv) | (KLT_EIG_PSD if psd else 0)
                                     | (KLT_EIG_WARM if psd and warm else 0)))
        return w, v

    def eig_topk_batch(self, a, k: int, psd: bool = False):
        """Top-k eigenpairs of each Hermitian matrix of a [count, n, n, 2] in one
        call: (w [count, k], vecs [count, k, n, 2]) on the same side as a."""
        count, n = int(a.shape[0]), int(a.shape[1])
        k = min(k, n)
        if _is_torch(a):
            import torch
            w = torch.empty((count, k), dtype=torch.float64, device=a.device)
            v = torch.empty((count, k, n, 2), dtype=torch.float64, device=a.device)
        else:
            w = np.empty((count, k), np.float64)
            v = np.empty((count, k, n, 2), np.float64)
        pa, ha = _ptr(a, np.float64)
        pw, hw = _ptr(w, np.float64)
        pv, hv = _ptr(v, np.float64)
        with _TorchOrder(self._h, a):
            check(lib().klt_eig_topk_batch(self._h, pa, count, n, k, pw, pv,
                                           _flags(ha, hw, hv) | (KLT_EIG_PSD if psd else 0)))
        return w, v

    def calc_kernels(self, tcc, nk: int, psd: bool = True, warm: bool = False):
        """calcKernels: (krn_r [nk, T], krn_i [nk, T], scales [nk]) float32; psd:
        the TCC is positive semidefinite (KLT_EIG_PSD; true of every TCC of a
        non-negative source); warm: start from the previous solve's block
        (KLT_EIG_WARM)."""
        T = int(tcc.shape[0])
        nk = min(nk, T)
        if _is_torch(tcc):
            import torch
            kr = torch.empty((nk, T), dtype=torch.float32, device=tcc.device)
            ki = torch.empty_like(kr)
            sc = torch.empty(nk, dtype=torch.float32, device=tcc.device)
        else:
            kr = np.empty((nk, T), np.float32)
            ki = np.empty((nk, T), np.float32)
            sc = np.empty(nk, np.float32)
        pt, ht = _ptr(tcc, np.float64)
        p1, h1 = _ptr(kr, np.float32)
        p2, h2 = _ptr(ki, np.float32)
        p3, h3 = _ptr(sc, np.float32)
        with _TorchOrder(self._h, tcc):
            check(lib().klt_calc_kernels(self._h, pt, T, nk, p1, p2, p3,
                                         _flags(ht, h1, h2, h3) | (KLT_EIG_PSD if psd else 0)
                                         | (KLT_EIG_WARM if psd and warm else 0)))
        return kr, ki, sc

    def hopkins_image(self, mask, tcc, Nx: int, Ny: int):
        """Hopkins image (calcImage + FT_c2r) from a TCC [T, T, 2]."""
        Ly, Lx = mask.shape
        if _is_torch(mask):
            import torch
            img = torch.empty_like(mask)
        else:
            img = np.empty((Ly, Lx), np.float32)
        pm, hm = _ptr(mask, np.float32)
        pt, ht = _ptr(tcc, np.float64)
        pi, hi = _ptr(img, np.float32)
        with _TorchOrder(self._h, mask, tcc):
            check(lib().klt_hopkins_image(self._h, pm, Lx, Ly, pt, Nx, Ny, pi,
                                          _flags(hm, ht, hi)))
        return img

    def resist_levels(self, images, vmin: float = 0.0, vmax: float = TARGET_INTENSITY):
        """Per-pixel class 0..6 of the reference's 7-level view (6 = I >= vmax
        when vmin = 0: the printed resist), uint8 of the images' shape."""
        shape = tuple(images.shape)
        Ly, Lx = shape[-2:]
        n = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
        if _is_torch(images):
            import torch
            out = torch.empty(shape, dtype=torch.uint8, device=images.device)
            po, ho = C.c_void_p(out.data_ptr()), False
        else:
            out = np.empty(shape, np.uint8)
            po, ho = out.ctypes.data_as(C.c_void_p), True
        pi, hi = _ptr(images, np.float32)
        with _TorchOrder(self._h, images):
            check(lib().klt_resist_levels(self._h, pi, n, Lx, Ly, vmin, vmax, po,
                                          _flags(hi, ho)))
        return out

    def kernels(self, kr, ki, scales, size_x: int, size_y: int) -> "KernelBank":
        return KernelBank(self, kr, ki, scales, size_x, size_y)

    def socs_forward(self, bank: "KernelBank", masks, images=None, plain: bool = False):
        """SOCS images of masks [B, Ly, Lx] (or [Ly, Lx]).

        plain=True runs klt_socs_image (the FP32 single-image path) per mask."""
        shape = tuple(masks.shape)
        Ly, Lx = shape[-2:]
        B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
        if images is None:
            if _is_torch(masks):
                import torch
                images = torch.empty_like(masks)
            else:
                images = np.empty(shape, np.float32)
        pm, hm = _ptr(masks, np.float32)
        pi, hi = _ptr(images, np.float32)
        fl = _flags(hm, hi)
        with _TorchOrder(self._h, masks, images):
            if plain:
                esz = 4 * Lx * Ly
                for i in range(B):
                    check(lib().klt_socs_image(self._h, bank.handle,
                                               C.c_void_p(pm.value + i * esz), Lx, Ly,
                                               C.c_void_p(pi.value + i * esz), fl))
            else:
                check(lib().klt_socs_forward(self._h, bank.handle, pm, B, Lx, Ly, pi, fl))
        return images

    def socs_forward_batch(self, banks, masks, images=None):
        """SOCS images of masks of different sizes under several kernel sets
        in one klt_socs_forward_batch call.  masks: list of [Ly, Lx] device
        (torch) tensors; banks: list of KernelBank.  Returns the list of
        images, images[m * len(banks) + s] = mask m under bank s."""
        import torch
        nm, nb = len(masks), len(banks)
        if images is None:
            images = [torch.empty_like(m) for m in masks for _ in banks]
        mp = (C.c_void_p * nm)(*[m.data_ptr() for m in masks])
        lx = (C.c_int * nm)(*[int(m.shape[-1]) for m in masks])
        ly = (C.c_int * nm)(*[int(m.shape[-2]) for m in masks])
        bp = (C.c_void_p * nb)(*[b.handle.value for b in banks])
        ip = (C.c_void_p * (nm * nb))(*[im.data_ptr() for im in images])
        with _TorchOrder(self._h, *masks):
            check(lib().klt_socs_forward_batch(self._h, nm, mp, lx, ly, nb, bp, ip, 0))
        return images

    def batch_latency(self) -> np.ndarray:
        """Per-image completion times (us from the batch start) of the last
        socs_forward_batch call."""
        n = lib().klt_socs_batch_latency(self._h, None, 0)
        if n < 0:
            check(n)
        out = np.empty(n, np.float32)
        rc = lib().klt_socs_batch_latency(self._h, out.ctypes.data_as(C.c_void_p), n)
        if rc < 0:
            check(rc)
        return out


def _empty_like(a, shape=None, dtype=np.float32):
    shape = tuple(a.shape) if shape is None else shape
    if _is_torch(a):
        import torch
        tdt = {np.float32: torch.float32, np.float64: torch.float64}[dtype]
        return torch.empty(shape, dtype=tdt, device=a.device)
    return np.empty(shape, dtype)


def socs_backward(ctx: "Context", bank: "KernelBank", masks, targets, a: float = 50.0,
                  thr: float = 0.25, want_images: bool = False):
    """(losses [B], grads [B, Ly, Lx], images or None) of the resist loss."""
    shape = tuple(masks.shape)
    Ly, Lx = shape[-2:]
    B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    grads = _empty_like(masks)
    losses = _empty_like(masks, (B,))
    images = _empty_like(masks) if want_images else None
    pm, hm = _ptr(masks, np.float32)
    pt, ht = _ptr(targets, np.float32)
    pg, hg = _ptr(grads, np.float32)
    pl, hl = _ptr(losses, np.float32)
    hs = [hm, ht, hg, hl]
    pi = None
    if images is not None:
        pi, hi = _ptr(images, np.float32)
        hs.append(hi)
    with _TorchOrder(ctx.handle, masks, targets):
        check(lib().klt_socs_backward(ctx.handle, bank.handle, pm, pt, B, Lx, Ly, a, thr, pg,
                                      pl, pi, _flags(*hs)))
    return losses, grads, images


def socs_vjp(ctx: "Context", bank: "KernelBank", masks, dl_di):
    """dL/dm [B, Ly, Lx] of the SOCS images of masks for a given dL/dI [B, Ly, Lx]
    (klt_socs_vjp)."""
    shape = tuple(masks.shape)
    Ly, Lx = shape[-2:]
    B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    grads = _empty_like(masks)
    pm, hm = _ptr(masks, np.float32)
    pd, hd = _ptr(dl_di, np.float32)
    pg, hg = _ptr(grads, np.float32)
    with _TorchOrder(ctx.handle, masks, dl_di):
        check(lib().klt_socs_vjp(ctx.handle, bank.handle, pm, pd, B, Lx, Ly, pg,
                                 _flags(hm, hd, hg)))
    return grads


def abbe_source_grad(ctx: "Context", masks, targets, src, NA: float, defocus: float, Nx: int,
                     Ny: int, a: float = 50.0, thr: float = 0.25, nmed: float = 1.0,
                     lam: float = 193.0, want_images: bool = False):
    """(losses [B], grad_src [B, n, n], Abbe images or None) of the resist loss."""
    shape = tuple(masks.shape)
    Ly, Lx = shape[-2:]
    B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    n = int(src.shape[0])
    grads = _empty_like(masks, (B, n, n))
    losses = _empty_like(masks, (B,))
    images = _empty_like(masks) if want_images else None
    pm, hm = _ptr(masks, np.float32)
    pt, ht = _ptr(targets, np.float32)
    ps, hs_ = _ptr(src, np.float32)
    pg, hg = _ptr(grads, np.float32)
    pl, hl = _ptr(losses, np.float32)
    hs = [hm, ht, hs_, hg, hl]
    pi = None
    if images is not None:
        if not _is_torch(images):
            raise TypeError("Abbe images are returned on the device only")
        pi, hi = _ptr(images, np.float32)
        hs.append(hi)
    with _TorchOrder(ctx.handle, masks, targets, src):
        check(lib().klt_abbe_source_grad(ctx.handle, pm, pt, B, Lx, Ly, ps, n, lam, NA,
                                         defocus, nmed, Nx, Ny, a, thr, pg, pl, pi, _flags(*hs)))
    return losses, grads, images


class KernelBank:
    def __init__(self, ctx: Context, kr, ki, scales, size_x: int, size_y: int):
        self.ctx = ctx
        self._h = C.c_void_p()
        pr, h1 = _ptr(kr, np.float32)
        pim, h2 = _ptr(ki, np.float32)
        ps, h3 = _ptr(scales, np.float32)
        nk = int(np.prod(scales.shape))
        with _TorchOrder(ctx.handle, kr, ki, scales):
            check(lib().klt_kernels_create(ctx.handle, pr, pim, ps, nk, size_x, size_y,
                                           _flags(h1, h2, h3), C.byref(self._h)))
        self.nk, self.size_x, self.size_y = nk, size_x, size_y

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().klt_kernels_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smo_run(ctx: "Context", masks, targets, src: np.ndarray, NA: float, steps: int,
            step_mask: float = 0.05, step_src: float = 0.05, nk: int = 24, defocus: float = 0.0,
            nmed: float = 1.0, a: float = 50.0, thr: float = 0.25, lam: float = 193.0):
    """klt_smo_run: masks (torch GPU tensor [B, Ly, Lx], updated in place) and
    src (numpy [n, n] float32, updated in place); returns (loss history
    [steps + 1], printed-resist mismatch counts [steps + 1])."""
    if not _is_torch(masks) or not _is_torch(targets):
        raise TypeError("smo_run takes device (torch) masks and targets")
    if not (isinstance(src, np.ndarray) and src.dtype == np.float32 and src.flags.c_contiguous):
        raise TypeError("src must be a C-contiguous float32 numpy array")
    shape = tuple(masks.shape)
    Ly, Lx = shape[-2:]
    B = int(np.prod(shape[:-2])) if len(shape) > 2 else 1
    hist = np.zeros(steps + 1, np.float64)
    mis = np.zeros(steps + 1, np.int64)
    pm, _ = _ptr(masks, np.float32)
    pt, _ = _ptr(targets, np.float32)
    with _TorchOrder(ctx.handle, masks, targets):
        check(lib().klt_smo_run(ctx.handle, pm, pt, B, Lx, Ly, src.ctypes.data_as(C.c_void_p),
                                int(src.shape[0]), lam, NA, defocus, nmed, nk, a, thr, step_mask,
                                step_src, steps, hist.ctypes.data_as(C.c_void_p),
                                mis.ctypes.data_as(C.c_void_p)))
    return hist, mis
