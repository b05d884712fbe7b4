"""FP64 eigensolver on the GPU (klt_eig_topk / klt_calc_kernels) against the
oracle's Householder + QL solver (itself pinned to LAPACK zheevr):

* eigenvalues within 1e-12 relative of the oracle;
* residual ||A v - lambda v|| <= 1e-11 ||A||, unit norm, orthogonality;
* eigenvector phases are arbitrary and degenerate pairs are common (an
  annular source gives exact pairs), so vectors are compared through cluster
  projectors sum_{q in cluster} v_q v_q^H;
* calcKernels: reversal krn[j] = z[T-1-j], float storage, and the SOCS image
  from the GPU kernels equals the image from the oracle's kernels.
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


def _check(a, w, v, wo, vo, k):
    nrm = np.abs(np.linalg.eigvalsh(a)).max()
    assert np.abs(w - wo).max() < 1e-12 * nrm
    for q in range(k):
        assert np.linalg.norm(a @ v[q] - w[q] * v[q]) < 1e-11 * nrm
    G = v @ v.conj().T
    assert np.abs(G - np.eye(k)).max() < 1e-11
    for (i, j) in _clusters(wo, 1e-9 * nrm):
        if j == k and j < a.shape[0]:
            continue  # cluster cut by k: its projector is not defined by k vectors
        P1 = v[i:j].T @ v[i:j].conj()
        P2 = vo[i:j].T @ vo[i:j].conj()
        assert np.abs(P1 - P2).max() < 1e-9


def _tcc(L, kind="Annular", NA=1.35, df=0.0, nmed=1.44, **kw):
    src = O.source(kind, kw.pop("n", 51), **kw)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(NA, L, L, sig)
    return O.tcc(src, sig, NA, df, L, L, Nx, Ny, nmed=nmed), Nx, Ny


@pytest.mark.parametrize("case", ["cfg0", "cfg1", "dipole_defocus"])
def test_eig_tcc_matches_oracle(ctx, case):
    print(f"case eig {case}", flush=True)
    if case == "cfg0":
        a, _, _ = _tcc(512, inner=0.6, outer=0.9)
        k = 24
    elif case == "cfg1":
        a, _, _ = _tcc(W.PERIOD_29, inner=0.6, outer=0.9)
        k = 24
    else:
        a, _, _ = _tcc(900, "Dipole", NA=1.35, df=0.2361, n=101, radius=0.2, offset=0.7, onx=1)
        k = 32
    wo, vo = O.eigh_topk(a, k)
    w, v = ctx.eig_topk(_pack(a), k)
    _check(a, w, _c(v), wo, vo, k)


@pytest.mark.parametrize("n,k,seed", [(1, 1, 0), (2, 2, 1), (3, 3, 2), (40, 40, 3), (257, 17, 4)])
def test_eig_random_hermitian_with_degeneracies(ctx, n, k, seed):
    print(f"case eig random n={n}", flush=True)
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.random(n))[::-1] * 10
    if n >= 6:
        lam[1] = lam[2] = lam[1]      # exact pair
        lam[3] = lam[4] = lam[5] = lam[3]  # exact triple
    a = (Q * lam) @ Q.conj().T
    a = 0.5 * (a + a.conj().T)
    wo, vo = O.eigh_topk(a, k)
    w, v = ctx.eig_topk(_pack(a), k)
    _check(a, w, _c(v), wo, vo, k)


def _reflected(n, lam, rng, nref=3):
    """Q diag(lam) Q^H with Q = H_nref ... H_1 a product of random Householder
    reflectors H = I - 2 u u^H, built in O(n^2); its eigenvectors are Q e_i."""
    a = np.diag(lam.astype(np.complex128))
    us = []
    for _ in range(nref):
        u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        u /= np.linalg.norm(u)
        us.append(u)
        au, uha = a @ u, u.conj() @ a
        uau = u.conj() @ au
        a = a - 2 * np.outer(u, uha) - 2 * np.outer(au, u.conj()) + 4 * uau * np.outer(u, u.conj())
    a = 0.5 * (a + a.conj().T)

    def vec(i):
        v = np.zeros(n, np.complex128)
        v[i] = 1.0
        for u in us:
            v = v - 2 * u * (u.conj() @ v)
        return v
    return a, vec


# Sizes past the other tests' paths: n = 1500 and 2100 take the
# back-transform's 8-column register path and its LDS path, n = 3500 the
# two-launch Householder steps (rows too long for the fused step's LDS).  The
# matrices are built with known eigenpairs (no O(n^3) reference solve).
@pytest.mark.parametrize("n,k,seed", [(1500, 12, 5), (2100, 8, 6), (3500, 6, 7)])
def test_eig_large_known_spectrum(ctx, n, k, seed):
    print(f"case eig known spectrum n={n}", flush=True)
    rng = np.random.default_rng(seed)
    lam = rng.random(n) * 10
    order = np.argsort(lam)[::-1]
    lam[order[2]] = lam[order[1]]                        # exact pair
    lam[order[4]] = lam[order[5]] = lam[order[3]]        # exact triple
    order = np.argsort(lam, kind="stable")[::-1]
    a, vec = _reflected(n, lam, rng)
    w, v = ctx.eig_topk(_pack(a), k)
    v = _c(v)
    nrm = np.abs(lam).max()
    assert np.abs(w - lam[order[:k]]).max() < 1e-12 * nrm
    for q in range(k):
        assert np.linalg.norm(a @ v[q] - w[q] * v[q]) < 1e-11 * nrm
    assert np.abs(v @ v.conj().T - np.eye(k)).max() < 1e-11
    # the top eigenvalue is simple: its vector matches Q e_i up to phase; the
    # pair and the triple through their projectors
    for (i, j) in [(0, 1), (1, 3), (3, 6)]:
        if j > k:
            continue
        P1 = v[i:j].T @ v[i:j].conj()
        ref = np.stack([vec(order[q]) for q in range(i, j)])
        P2 = ref.T @ ref.conj()
        assert np.abs(P1 - P2).max() < 1e-9


def test_eig_batch_equals_single_calls(ctx):
    """klt_eig_topk_batch: each matrix's eigenpairs bit-identical to its own
    klt_eig_topk call (the same kernels over a grid with one more dimension),
    for three TCCs of T = 169 (annular source at three defocus values) and a
    random Hermitian matrix with an exact pair and triple."""
    print("case eig batch of 4", flush=True)
    mats = [_tcc(512, inner=0.6, outer=0.9, df=df)[0] for df in (0.0, 0.15, -0.3)]
    n = mats[0].shape[0]
    rng = np.random.default_rng(11)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.random(n))[::-1] * 10
    lam[1] = lam[2] = lam[1]
    lam[3] = lam[4] = lam[5] = lam[3]
    r = (Q * lam) @ Q.conj().T
    mats.append(0.5 * (r + r.conj().T))
    k = 24
    packed = np.stack([_pack(a) for a in mats])
    wb, vb = ctx.eig_topk_batch(packed, k)
    assert wb.shape == (4, k) and vb.shape == (4, k, n, 2)
    for i, a in enumerate(mats):
        w, v = ctx.eig_topk(_pack(a), k)
        assert np.array_equal(wb[i], w), i
        assert np.array_equal(vb[i], v), i
    wo, vo = O.eigh_topk(mats[1], k)
    _check(mats[1], wb[1], _c(vb[1]), wo, vo, k)


@pytest.mark.parametrize("n,count", [(3, 2), (40, 3), (64, 2), (65, 2), (67, 1), (97, 2)])
def test_eig_small_orders_and_tail(ctx, n, count):
    """Orders around the last-columns block (the final 64 columns run in one
    register-resident block per matrix): n = 3, 40 and 64 run entirely there,
    65, 67 and 97 take 1, 3 and 33 device-wide steps first.  Batched results
    equal single calls bit for bit and match the oracle."""
    print(f"case eig small n={n} count={count}", flush=True)
    rng = np.random.default_rng(20 + n)
    mats = []
    for _ in range(count):
        a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        mats.append(0.5 * (a + a.conj().T))
    k = min(12, n)
    wb, vb = ctx.eig_topk_batch(np.stack([_pack(a) for a in mats]), k)
    for i, a in enumerate(mats):
        w, v = ctx.eig_topk(_pack(a), k)
        assert np.array_equal(wb[i], w) and np.array_equal(vb[i], v), i
        wo, vo = O.eigh_topk(a, k)
        _check(a, w, _c(v), wo, vo, k)


def test_calc_kernels_layout_and_image(ctx):
    print("case calc_kernels", flush=True)
    a, Nx, Ny = _tcc(512, inner=0.6, outer=0.9)
    kr, ki, sc = ctx.calc_kernels(_pack(a), 24)
    wo, vo = O.eigh_topk(a, 24)
    kro, kio, sco = O.kernels_from_eig(vo, wo)
    assert np.abs(sc - sco).max() <= 2e-7 * sco[0]
    # kernels equal up to phase inside each cluster: compare images
    m = W.manhattan(512, 3)
    assert relerr(O.socs_image(m, kr, ki, sc, Nx, Ny), O.socs_image(m, kro, kio, sco, Nx, Ny)) < 1e-6
    T = a.shape[0]
    z = kr[0].astype(np.float64) + 1j * ki[0]
    v0 = z[::-1]  # undo the reversal: eigenvector of the largest eigenvalue
    assert np.linalg.norm(a @ v0 - wo[0] * v0) < 1e-5 * wo[0]
    assert abs(np.linalg.norm(v0) - 1) < 1e-6
    # nk > T is clamped like the reference
    kr2, _, sc2 = ctx.calc_kernels(_pack(a[:9, :9]), 20)
    assert kr2.shape == (9, 9) and sc2.shape == (9,)
    assert T == 169


@pytest.mark.parametrize("kind", ["zero", "diagonal", "diagonal_repeated"])
def test_eig_zero_and_diagonal_inputs(ctx, kind):
    """Tridiagonals with zero off-diagonals: an all-zero matrix (every pivot of
    the inverse iteration's LU is zero), distinct and repeated diagonal
    entries.  Finite results, exact eigenvalues, orthonormal eigenvectors."""
    print(f"case eig {kind}", flush=True)
    n, k = 96, 12
    rng = np.random.default_rng(5)
    if kind == "zero":
        d = np.zeros(n)
    elif kind == "diagonal":
        d = rng.permutation(n).astype(np.float64) - 20.0
    else:
        d = np.repeat(np.arange(n // 4, dtype=np.float64), 4)
        rng.shuffle(d)
    a = np.diag(d).astype(np.complex128)
    w, v = ctx.eig_topk(_pack(a), k)
    v = _c(v)
    assert np.all(np.isfinite(w)) and np.all(np.isfinite(v))
    wo = np.sort(d)[::-1][:k]
    scale = max(1.0, np.abs(d).max())
    assert np.abs(w - wo).max() < 1e-12 * scale
    for q in range(k):
        assert np.linalg.norm(a @ v[q] - w[q] * v[q]) < 1e-11 * scale
    assert np.abs(v @ v.conj().T - np.eye(k)).max() < 1e-11
