"""Index-for-index numpy restatement of K-Litho's FFT imaging pipeline.

TEST INFRASTRUCTURE ONLY (see oracle.py).  The C oracle evaluates the SOCS and
Hopkins images through their closed forms; this module follows the
reference's own sequence of shifts, FFTW-convention transforms and
zero-padding, so that the tests can pin the closed form (centres, signs,
scaling, odd sizes) against the procedure the reference actually runs.

numpy's FFT conventions map onto FFTW's as: fft2 = FFTW_FORWARD,
L*ifft2 = FFTW_BACKWARD (unnormalised), rfft2 = r2c, (Lx*Ly)*irfft2 = c2r.
"""
from __future__ import annotations

import numpy as np


def my_shift(a: np.ndarray, type_x: bool, type_y: bool) -> np.ndarray:
    """myShift (klitho_socs.cpp:55-68): out[(y+yh)%Y][(x+xh)%X] = in[y][x]."""
    Y, X = a.shape
    xh = X // 2 if type_x else (X + 1) // 2
    yh = Y // 2 if type_y else (Y + 1) // 2
    return np.roll(a, (yh, xh), axis=(0, 1))


def ft_r2c(m: np.ndarray) -> np.ndarray:
    """FT_r2c (klitho_socs.cpp:71-121): centred full spectrum, / (Lx Ly).

    Returns out[y][x] with the reference's layout (frequency (x - Lx/2,
    y - Ly/2) at index (y, x) after the shuffles)."""
    Ly, Lx = m.shape
    d = my_shift(m.astype(np.float64), False, False)
    datac = np.fft.rfft2(d)  # Ly x (Lx/2+1)
    wt = Lx // 2 + 1
    AY = (Ly + 1) // 2
    BY = Ly - AY
    R = np.zeros((Ly, wt), complex)
    R[:BY] = datac[AY:AY + BY]
    R[BY:] = datac[:Ly - BY]
    Lm = np.zeros((Ly, wt), complex)
    if Ly % 2 == 0:
        Lm[0, :] = np.conj(R[0, ::-1])
        for y in range(1, Ly):
            Lm[y, :] = np.conj(R[Ly - y, ::-1])
    else:
        for y in range(Ly):
            Lm[y, :] = np.conj(R[Ly - 1 - y, ::-1])
    out = np.zeros((Ly, Lx), complex)
    out[:, :wt] = Lm
    sxh = Lx // 2
    out[:, sxh:] = R[:, :Lx - sxh]
    return out / (Lx * Ly)


def fi(img_c: np.ndarray, Lx: int, Ly: int) -> np.ndarray:
    """FI (klitho_socs.cpp:159-194): exact Fourier interpolation."""
    Sy, Sx = img_c.shape
    if Lx - Sx < 0 or Ly - Sy < 0:
        raise ValueError("ERROR in FI()")
    difY = Ly - Sy
    d = my_shift(img_c, False, False)
    spec = np.fft.rfft2(d) / (Sx * Sy)
    out = np.zeros((Ly, Lx // 2 + 1), complex)
    hx = Sx // 2 + 1
    out[:Sy // 2 + 1, :hx] = spec[:Sy // 2 + 1, :hx]
    out[Sy // 2 + 1 + difY:Sy + difY, :hx] = spec[Sy // 2 + 1:Sy, :hx]
    r = np.fft.irfft2(out, s=(Ly, Lx)) * (Lx * Ly)
    return my_shift(r, False, False)


def calc_socs(mskf: np.ndarray, krns: np.ndarray, scales: np.ndarray, Nx: int,
              Ny: int) -> np.ndarray:
    """calcSOCS (klitho_socs.cpp:196-241); krns: nk x T complex."""
    Ly, Lx = mskf.shape
    sx, sy = 4 * Nx + 1, 4 * Ny + 1
    difX, difY = sx - (2 * Nx + 1), sy - (2 * Ny + 1)
    Lxh, Lyh = Lx // 2, Ly // 2
    tmp = np.zeros((sy, sx))
    band = mskf[Lyh - Ny:Lyh + Ny + 1, Lxh - Nx:Lxh + Nx + 1]
    for k in range(krns.shape[0]):
        bwd = np.zeros((sy, sx), complex)
        bwd[difY:difY + 2 * Ny + 1, difX:difX + 2 * Nx + 1] = (
            krns[k].reshape(2 * Ny + 1, 2 * Nx + 1) * band)
        out = np.fft.ifft2(bwd) * (sx * sy)
        tmp += float(scales[k]) * (out.real ** 2 + out.imag ** 2)
    tmpp = my_shift(tmp, True, True)
    return fi(tmpp, Lx, Ly)


def ft_c2r(spec: np.ndarray) -> np.ndarray:
    """FT_c2r (klitho_tcc.cpp:151-183) without its out-of-bounds store."""
    Ly, Lx = spec.shape
    wt = Lx // 2 + 1
    xh, yh = (Lx + 1) // 2, (Ly + 1) // 2
    datac = np.zeros((Ly, wt), complex)
    for y in range(Ly):
        for x in range(xh - 1, Lx):
            sxx = (x + xh) % Lx
            syy = (y + yh) % Ly
            if sxx < wt:
                datac[syy, sxx] = spec[y, x]
    if Lx % 2 == 0:
        R = np.zeros(Ly, complex)
        R[0] = np.conj(spec[0, 0])
        for y in range(1, Ly):
            R[y] = np.conj(spec[Ly - y, 0])
        datac[:, wt - 1] = R
    r = np.fft.irfft2(datac, s=(Ly, Lx)) * (Lx * Ly)
    return my_shift(r, False, False)


def calc_image(mskf: np.ndarray, tcc: np.ndarray, Nx: int, Ny: int) -> np.ndarray:
    """calcImage (klitho_tcc.cpp:270-290): Hopkins spectrum, reference layout."""
    Ly, Lx = mskf.shape
    T = tcc.shape[0]
    half = (T - 1) // 2
    Lxh, Lyh = Lx // 2, Ly // 2
    W = 2 * Nx + 1
    imgf = np.zeros((Ly, Lx), complex)
    for ny2 in range(-2 * Ny, 2 * Ny + 1):
        for nx2 in range(0, 2 * Nx + 1):
            v = 0j
            for ny1 in range(-Ny, Ny + 1):
                for nx1 in range(-Nx, Nx + 1):
                    if abs(nx2 + nx1) <= Nx and abs(ny2 + ny1) <= Ny:
                        v += (mskf[ny2 + ny1 + Lyh, nx2 + nx1 + Lxh]
                              * np.conj(mskf[ny1 + Lyh, nx1 + Lxh])
                              * tcc[ny1 * W + nx1 + half, (ny2 + ny1) * W + nx2 + nx1 + half])
            imgf[ny2 + Lyh, nx2 + Lxh] = v
    return imgf


def socs_pipeline(mask: np.ndarray, kr: np.ndarray, ki: np.ndarray, scales: np.ndarray,
                  Nx: int, Ny: int) -> np.ndarray:
    """klitho_socs end to end: FT_r2c then calcSOCS (klitho_socs.cpp:338-343)."""
    krns = kr.astype(np.float64) + 1j * ki.astype(np.float64)
    return calc_socs(ft_r2c(mask), krns, scales.astype(np.float64), Nx, Ny)


def hopkins_pipeline(mask: np.ndarray, tcc: np.ndarray, Nx: int, Ny: int) -> np.ndarray:
    """klitho_tcc image.bin path: FT_r2c, calcImage, FT_c2r (:510,534,537)."""
    return ft_c2r(calc_image(ft_r2c(mask), tcc, Nx, Ny))
