"""The CPU baseline's restatement of the reference path (oracle/cpu_ref_path.cpp:
FT_r2c + calcSOCS + FI on FFTW3, calcKernels' zheevr_) against the FP64 oracle
and the numpy port of the same FFTW procedure, so that the baseline bench.py
times computes the reference's images:

* images within relative L2 1e-7 of the oracle (the output is float32, as the
  reference's) at even, odd and unequal sizes, and of the numpy port;
* several worker threads give the same images as one;
* the zheevr_ eigenpairs: eigenvalues within 1e-12 of the oracle's (relative
  to the largest), projectors of the eigenvectors within 1e-9;
* 4N+1 > L is refused (the reference's "ERROR in FI()").
Skipped when no library exporting the FFTW3 interface and zheevr_ loads.
"""
import numpy as np
import pytest

import cpu_ref as CR
import oracle as O
import ref_pipeline as RP
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.skipif(not CR.available(), reason=CR.reason())


def _bank(rng, Nx, Ny, nk):
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = rng.standard_normal((nk, T)).astype(np.float32)
    ki = rng.standard_normal((nk, T)).astype(np.float32)
    sc = np.sort(rng.random(nk).astype(np.float32))[::-1].copy()
    return kr, ki, sc


@pytest.mark.parametrize("Lx,Ly,Nx,Ny", [(64, 64, 3, 3), (63, 65, 3, 2), (96, 80, 5, 6),
                                         (512, 512, 14, 14), (255, 256, 9, 7)])
def test_cpu_ref_images_vs_oracle(Lx, Ly, Nx, Ny):
    rng = np.random.default_rng(Lx + 3 * Ly)
    kr, ki, sc = _bank(rng, Nx, Ny, 4)
    m = (rng.random((2, Ly, Lx)) > 0.5).astype(np.float32)
    got = CR.socs_images(m, kr, ki, sc, Nx, Ny)
    for i in range(2):
        ref = O.socs_image(m[i], kr, ki, sc, Nx, Ny)
        assert relerr(got[i], ref) < 1e-7
    if Lx * Ly <= 96 * 96:
        assert relerr(got[0], RP.socs_pipeline(m[0], kr, ki, sc, Nx, Ny)) < 1e-7


def test_cpu_ref_threads_give_the_same_images():
    rng = np.random.default_rng(7)
    kr, ki, sc = _bank(rng, 6, 6, 3)
    m = np.stack([W.manhattan(256, 10 + i) for i in range(5)])
    one = CR.socs_images(m, kr, ki, sc, 6, 6, threads=1)
    many = CR.socs_images(m, kr, ki, sc, 6, 6, threads=3)
    assert np.array_equal(one, many)


def test_cpu_ref_zheevr_vs_oracle():
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(1.35, 512, 512, sig)
    t = O.tcc(src, sig, 1.35, 0.0, 512, 512, Nx, Ny, nmed=1.44)
    w, v = CR.zheevr_topk(t, 24, threads=2)
    wo, vo = O.eigh_topk(t, 24)
    assert np.abs(w - wo).max() < 1e-12 * wo[0]
    i = 0
    while i < 24:  # projectors of (numerically) equal eigenvalues
        j = i + 1
        while j < 24 and abs(wo[j] - wo[i]) < 1e-9 * wo[0]:
            j += 1
        if j < 24:
            P1 = v[i:j].T @ v[i:j].conj()
            P2 = vo[i:j].T @ vo[i:j].conj()
            assert np.abs(P1 - P2).max() < 1e-9
        i = j


def test_cpu_ref_rejects_small_period():
    rng = np.random.default_rng(1)
    kr, ki, sc = _bank(rng, 4, 4, 1)
    with pytest.raises(ValueError):
        CR.socs_images(np.zeros((16, 16), np.float32), kr, ki, sc, 4, 4)
