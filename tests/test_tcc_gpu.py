"""TCC build on the GPU (klt_tcc_build) against the FP64 oracle.

* FP64 Gram: entries within 1e-12 of the oracle at zero defocus (unit pupil,
  same products, only the summation order differs) and within 1e-6 with
  defocus (device cosf/sinf vs glibc's differ by at most an ulp or two).
  The diagonal TCC[i][i] = sum_s S_s [i inside pupil(s)] pins every discrete
  pupil decision: one flipped sample moves it by S_s >= 1e-4.
* split-FP16 MFMA Gram: relative Frobenius error <= 1e-5 against the FP64
  oracle and the FP64 GPU Gram on complex entries; the imaginary part alone
  within 1e-5 of the norm (annular and dipole sources, with and without
  defocus).
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu


def _annular():
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    return src, O.outer_sigma(src)


def _cplx(a):
    return a[..., 0] + 1j * a[..., 1]


@pytest.mark.parametrize("L,NA,df,nmed", [(512, 1.35, 0.0, 1.44), (1088, 1.35, 0.0, 1.44),
                                          (700, 0.95, 0.25, 1.0), (512, 1.35, -0.2361, 1.44)])
def test_tcc_fp64_vs_oracle(ctx, L, NA, df, nmed):
    print(f"case tcc fp64 L={L} NA={NA} df={df}", flush=True)
    src, sig = _annular()
    Nx, Ny = O.sizing(NA, L, L, sig)
    ref = O.tcc(src, sig, NA, df, L, L, Nx, Ny, nmed=nmed)
    got = _cplx(ctx.tcc_build(src, sig, NA, df, L, L, Nx, Ny, precision=0, nmed=nmed))
    tol = 1e-12 if df == 0.0 else 1e-6
    assert np.abs(got - ref).max() < tol * np.abs(ref).max()
    # one flipped pupil sample would move a diagonal entry by S_s >= 1e-4
    assert np.abs(np.diag(got).real - np.diag(ref).real).max() < (1e-12 if df == 0 else 1e-6)
    assert np.abs(got - got.conj().T).max() == 0.0


SPLIT_CASES = [("Annular", 512, 1.35, 0.0, 1.44), ("Annular", 1088, 1.35, 0.0, 1.44),
               ("Annular", 1280, 1.35, 0.4722, 1.44), ("Dipole", 700, 0.95, -0.25, 1.0),
               ("Dipole", 1024, 1.35, 0.2361, 1.44)]


@pytest.mark.parametrize("kind,L,NA,df,nmed", SPLIT_CASES)
def test_tcc_split16_vs_oracle(ctx, kind, L, NA, df, nmed):
    """split-FP16 MFMA Gram against the FP64 oracle and the FP64 GPU Gram, on the
    complex entries.  With defocus the imaginary part (formed by the pair-swap
    operand X') carries a large share of the norm; it is checked on its own."""
    print(f"case tcc split16 {kind} L={L} NA={NA} df={df}", flush=True)
    if kind == "Annular":
        src, sig = _annular()
    else:
        src = O.source("Dipole", 101, radius=0.2, offset=0.7, onx=1)
        sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(NA, L, L, min(sig, 1.0))
    ref = O.tcc(src, sig, NA, df, L, L, Nx, Ny, nmed=nmed)
    a = _cplx(ctx.tcc_build(src, sig, NA, df, L, L, Nx, Ny, precision=0, nmed=nmed))
    b = _cplx(ctx.tcc_build(src, sig, NA, df, L, L, Nx, Ny, precision=1, nmed=nmed))
    assert relerr(b, ref) < 1e-5
    assert relerr(b, a) < 1e-5
    nref = np.linalg.norm(ref)
    assert np.linalg.norm(b.imag - ref.imag) < 1e-5 * nref
    if df != 0.0:
        assert np.linalg.norm(ref.imag) > 0.1 * nref
    assert np.abs(b - b.conj().T).max() == 0.0


def test_tcc_device_pointers_and_dipole(ctx):
    import torch
    print("case tcc dipole device", flush=True)
    src = O.source("Dipole", 101, radius=0.2, offset=0.7, onx=1)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(0.9, 600, 600, sig)
    ref = O.tcc(src, sig, 0.9, 0.1, 600, 600, Nx, Ny)
    dsrc = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    got = ctx.tcc_build(dsrc, sig, 0.9, 0.1, 600, 600, Nx, Ny, precision=0)
    torch.cuda.synchronize()
    got = _cplx(got.cpu().numpy())
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()


def test_tcc_rejects_na_above_medium(ctx):
    src, sig = _annular()
    with pytest.raises(klt.KltError) as e:
        ctx.tcc_build(src, sig, 1.35, 0.3, 512, 512, 6, 6, precision=0, nmed=1.0)
    assert "NA" in e.value.msg


@pytest.mark.parametrize("L,NA", [(256, 0.9), (255, 0.9), (512, 1.35)])
def test_hopkins_image_vs_oracle(ctx, L, NA):
    """klt_hopkins_image (calcImage + FT_c2r) against the oracle's Hopkins image."""
    print(f"case hopkins L={L}", flush=True)
    src, sig = _annular()
    Nx, Ny = O.sizing(NA, L, L, sig)
    t = O.tcc(src, sig, NA, 0.0, L, L, Nx, Ny, nmed=1.44)
    m = W.manhattan(L, 42)
    got = ctx.hopkins_image(m, np.ascontiguousarray(np.stack([t.real, t.imag], -1)), Nx, Ny)
    assert relerr(got, O.hopkins_image(m, t, Nx, Ny)) < 1e-5
