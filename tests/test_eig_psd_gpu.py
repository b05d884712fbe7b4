"""The positive-semidefinite eigen path (KLT_EIG_PSD: block subspace iteration,
X <- cholqr(A^2 X), with Rayleigh-Ritz, eig_subspace.hip) against the FP64
oracle's solver (pinned to LAPACK zheevr), on the TCCs the reference
eigen-decomposes (K-Litho-TCC/klitho_tcc.cpp:301-377):

* eigenvalues within 1e-12 ||A||, residuals within 1e-11 ||A||,
  orthonormality to 1e-11, and projectors of clusters (eigenvalues closer
  than 1e-6 ||A||) within 1e-8 (phases are arbitrary), at T = 169, 841 and
  1225 (annular, blurred dipole with and without defocus);
* the kernels the subspace path gives image a mask like the oracle's;
* a batch equals its single calls bit for bit;
* the Householder path takes over, with correct results, for k above the
  block's reach and for a flat spectrum the filter cannot separate; the
  kernel names recorded by the profiler show which path ran;
* KLT_EIG_WARM: after a small change of the source, a solve that starts from
  the previous solve's block meets the same checks in fewer A X products; one
  that starts from the block of an unrelated matrix of the same order meets
  them too.
"""
import numpy as np
import pytest

import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu


def _c(a):
    return a[..., 0] + 1j * a[..., 1]


def _pack(a):
    return np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)


def _clusters(w, tol):
    out, i = [], 0
    while i < len(w):
        j = i + 1
        while j < len(w) and abs(w[j] - w[i]) < tol:
            j += 1
        out.append((i, j))
        i = j
    return out


def _check(a, w, v, k):
    wo, vo = O.eigh_topk(a, k + 1)
    nrm = abs(wo[0])
    assert np.abs(w - wo[:k]).max() < 1e-12 * nrm
    for q in range(k):
        assert np.linalg.norm(a @ v[q] - w[q] * v[q]) < 1e-11 * nrm
    assert np.abs(v @ v.conj().T - np.eye(k)).max() < 1e-11
    for (s, e) in _clusters(wo[:k + 1], 1e-6 * nrm):
        if e > k:
            continue  # cluster cut by k
        P1 = v[s:e].T @ v[s:e].conj()
        P2 = vo[s:e].T @ vo[s:e].conj()
        assert np.abs(P1 - P2).max() < 1e-8, (s, e)


def _tcc(case):
    if case.startswith("dipole"):
        src = W.blurred(O.source("Dipole", 101, radius=0.2, offset=0.7, onx=1))
        sig, L = min(O.outer_sigma(src), 1.0), W.PERIOD_35
        df = 0.4722 if case.endswith("50nm") else 0.0
    else:
        src = O.source("Annular", 51, inner=0.6, outer=0.9)
        sig, df = O.outer_sigma(src), 0.0
        L = 512 if case == "annular512" else W.PERIOD_29
    Nx, Ny = O.sizing(1.35, L, L, sig)
    return O.tcc(src, sig, 1.35, df, L, L, Nx, Ny, nmed=1.44)


@pytest.mark.parametrize("case,k", [("annular512", 24), ("annular1088", 24), ("dipole0", 32),
                                    ("dipole50nm", 32)])
def test_psd_path_vs_oracle(ctx, case, k):
    print(f"case eig psd {case} k={k}", flush=True)
    a = _tcc(case)
    ctx.set_profiling(True)
    w, v = ctx.eig_topk(_pack(a), k, psd=True)
    rep = ctx.profile_report()
    ctx.set_profiling(False)
    assert "k_zgemm" in rep and "k_hh_step" not in rep, sorted(rep)
    _check(a, w, _c(v), k)


def test_psd_kernels_image(ctx):
    print("case eig psd kernels -> image", flush=True)
    a = _tcc("annular1088")
    kr, ki, sc = ctx.calc_kernels(_pack(a), 24, psd=True)
    wo, vo = O.eigh_topk(a, 24)
    okr, oki, osc = O.kernels_from_eig(vo, wo)
    m = W.manhattan(W.PERIOD_29, 77)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    img = ctx.socs_forward(bank, m)
    assert relerr(img, O.socs_image(m, okr, oki, osc, 14, 14)) < 1e-5
    bank.close()


def test_psd_batch_equals_single_calls(ctx):
    print("case eig psd batch of 3 == single calls", flush=True)
    mats = [_tcc("dipole0"), _tcc("dipole50nm"), _tcc("dipole0") * 0.5]
    a3 = np.ascontiguousarray(np.stack([_pack(m) for m in mats]))
    wb, vb = ctx.eig_topk_batch(a3, 32, psd=True)
    for i, m in enumerate(mats):
        w1, v1 = ctx.eig_topk(_pack(m), 32, psd=True)
        assert np.array_equal(wb[i], w1) and np.array_equal(vb[i], v1), i


@pytest.mark.parametrize("kind", ["k_above_block", "flat_spectrum"])
def test_psd_falls_back_to_householder(ctx, kind):
    print(f"case eig psd fallback {kind}", flush=True)
    rng = np.random.default_rng(17)
    if kind == "k_above_block":
        a, k = _tcc("annular512"), 60
    else:  # eigenvalues 1 + 1e-3 U(0, 1): no decay beyond k
        n, k = 300, 24
        q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        d = 1.0 + 1e-3 * rng.random(n)
        a = (q * d) @ q.conj().T
        a = 0.5 * (a + a.conj().T)
    ctx.set_profiling(True)
    w, v = ctx.eig_topk(_pack(a), k, psd=True)
    rep = ctx.profile_report()
    ctx.set_profiling(False)
    assert "k_hh_step" in rep or "k_hh_tail" in rep, sorted(rep)
    wo, vo = O.eigh_topk(a, k)
    nrm = abs(wo[0])
    assert np.abs(w - wo).max() < 1e-12 * nrm
    vv = _c(v)
    for q in range(k):
        assert np.linalg.norm(a @ vv[q] - w[q] * vv[q]) < 1e-11 * nrm


def test_psd_warm_start(ctx):
    print("case eig psd warm start (annular 1088, T = 841)", flush=True)
    rng = np.random.default_rng(23)
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(1.35, W.PERIOD_29, W.PERIOD_29, sig)
    a1 = O.tcc(src, sig, 1.35, 0.0, W.PERIOD_29, W.PERIOD_29, Nx, Ny, nmed=1.44)
    src2 = src.copy()
    on = np.flatnonzero(src2 > 0)
    pick = rng.choice(on, max(1, len(on) // 20), replace=False)
    src2.reshape(-1)[pick] *= 0.9
    a2 = O.tcc(src2, sig, 1.35, 0.0, W.PERIOD_29, W.PERIOD_29, Nx, Ny, nmed=1.44)
    n = a2.shape[0]

    def products(a, warm):
        ctx.set_profiling(True)
        w, v = ctx.eig_topk(_pack(a), 24, psd=True, warm=warm)
        rep = ctx.profile_report()
        ctx.set_profiling(False)
        return w, _c(v), rep["k_zgemm"][0]

    w, v, cold = products(a2, False)
    _check(a2, w, v, 24)
    products(a1, False)
    w, v, warm = products(a2, True)
    print(f"A X products: cold {cold}, warm {warm}", flush=True)
    assert warm < cold
    _check(a2, w, v, 24)
    # a block of an unrelated matrix of the same order (decaying spectrum)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    d = 10.0 ** -np.linspace(0, 8, n)
    b = (q * d) @ q.conj().T
    products(0.5 * (b + b.conj().T), False)
    w, v, _ = products(a2, True)
    _check(a2, w, v, 24)
