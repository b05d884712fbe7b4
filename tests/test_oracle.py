"""Pins the FP64 oracle before anything is compared against it.

Checks, per oracle function:
  * sources / masks / kernel_info: bit-exact against K-Litho's own
    source.cpp, mask.cpp, file_io.cpp (compiled by oracle/Makefile into
    oracle/_ref) and the shipped fixture K-Litho-TCC/source/src_test1_size101.bin
    (copied under tests/golden/);
  * SOCS / Hopkins images: against an index-for-index numpy restatement of the
    reference's FFT pipeline (FT_r2c, calcSOCS + FI, calcImage + FT_c2r) at even
    and odd sizes;
  * eigensolver: against scipy.linalg.eigh(driver="evr") (LAPACK zheevr, the
    routine the reference calls);
  * TCC: Hermitian PSD, TCC(0,0) = sum of source weights inside the pupil,
    trace formula, point source = rank one with lambda = #pupil points,
    centro-symmetry, Abbe image == Hopkins image;
  * adjoint: central finite differences and the exact identities
    <m, dL/dm> = 2 sum_x G I (I is quadratic in m) and
    sum_s S_s dL/dS_s = sum_x G I_A (I_A is linear in S).
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O
import ref_pipeline as RP
from conftest import relerr

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _ref_or_skip():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return R


# ---------------------------------------------------------------- sources
SOURCE_CASES = [
    ("Annular", 51, dict(inner=0.6, outer=0.9)),
    ("Annular", 101, dict(inner=0.6, outer=0.9)),
    ("Annular", 50, dict(inner=0.3, outer=0.7)),
    ("Dipole", 101, dict(radius=0.2, offset=0.7, onx=1)),
    ("Dipole", 101, dict(radius=0.1, offset=0.5, onx=0)),
    ("Dipole", 40, dict(radius=0.25, offset=0.6, onx=1)),
    ("CrossQuadrupole", 101, dict(radius=0.1, offset=0.5)),
    ("CrossQuadrupole", 64, dict(radius=0.2, offset=0.6)),
    ("Point", 101, dict(px=0.6, py=0.2)),
    ("Point", 51, dict(px=-0.4, py=-0.3)),
    ("Point", 50, dict(px=0.0, py=0.0)),
]


@pytest.mark.parametrize("kind,n,kw", SOURCE_CASES)
def test_sources_bit_exact_vs_reference(kind, n, kw):
    R = _ref_or_skip()
    mine = O.source(kind, n, **kw)
    ref = np.zeros(n * n, np.float32)
    R.ref_create_source(ref, n, kind.encode(), kw.get("inner", 0.6), kw.get("outer", 0.9),
                        kw.get("radius", 0.1), kw.get("offset", 0.5), kw.get("onx", 1),
                        kw.get("px", 0.6), kw.get("py", 0.2), b"")
    assert np.array_equal(mine.ravel().view(np.uint32), ref.view(np.uint32))


def test_source_counts_and_sigma():
    a = O.source("Annular", 51, inner=0.6, outer=0.9)
    assert int((a > 0).sum()) == 820
    assert O.outer_sigma(a) == np.float32(0.88)
    d = O.source("Dipole", 101, radius=0.2, offset=0.7, onx=1)
    assert int((d > 0).sum()) == 634
    assert abs(O.outer_sigma(d) - 0.9) < 1e-6


def test_import_source_fixture():
    """Shipped import source: annulus sigma 0.4..0.6, 1576 points of 1/1576."""
    raw = np.fromfile(os.path.join(GOLD, "src_test1_size101.bin"), np.float32)
    assert raw.size == 101 * 101
    nz = raw[raw != 0]
    assert nz.size == 1576
    assert np.allclose(nz, 1.0 / 1576, rtol=1e-6)
    R = O.ref()
    if R is not None:
        ref = np.zeros(101 * 101, np.float32)
        R.ref_create_source(ref, 101, b"Import", 0, 0, 0, 0, 0, 0, 0,
                            os.path.join(GOLD, "src_test1_size101.bin").encode())
        mine = raw.copy()
        O.lib().or_normalize(mine, mine.size)
        assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


# ---------------------------------------------------------------- masks
@pytest.mark.parametrize("L,ms,lw,sw,hz", [(512, 512, 32, 32, 0), (512, 300, 20, 44, 1),
                                          (65, 33, 3, 5, 0), (64, 63, 7, 2, 1)])
def test_linespace_mask_bit_exact_vs_reference(L, ms, lw, sw, hz):
    R = _ref_or_skip()
    inner = np.zeros(ms * ms, np.float32)
    O.lib().or_mask_linespace(inner, ms, ms, lw, sw, hz)
    mine = np.zeros(L * L, np.float32)
    O.lib().or_mask_place(inner, ms, ms, mine, L, L, 1.0)
    ref = np.zeros(L * L, np.float32)
    R.ref_create_mask(ref, L, L, ms, ms, b"LineSpace", lw, sw, hz, b"", 1.0)
    assert np.array_equal(mine, ref)


def test_reference_kernel_info_fixture(tmp_path):
    """Shipped kernel_info.txt: 33x33, 3 kernels, descending, degenerate pair."""
    R = _ref_or_skip()
    sx, sy = C.c_int(), C.c_int()
    sc = np.zeros(8, np.float32)
    nk = R.ref_read_kernel_info(os.path.join(GOLD, "kernel_info.txt").encode(),
                                C.byref(sx), C.byref(sy), sc, 8)
    assert (sx.value, sy.value, nk) == (33, 33, 3)
    assert np.allclose(sc[:3], [78.9658, 37.2823, 37.2823])
    for i in range(3):
        r = np.fromfile(os.path.join(GOLD, f"krn_{i}_r.bin"), np.float32)
        im = np.fromfile(os.path.join(GOLD, f"krn_{i}_i.bin"), np.float32)
        assert r.size == im.size == 33 * 33
        assert abs(float(np.sum(r.astype(np.float64) ** 2 + im.astype(np.float64) ** 2)) - 1) < 1e-5


# ---------------------------------------------------------------- sizing
def test_sizing_matches_configs():
    s = O.outer_sigma(O.source("Annular", 51, inner=0.6, outer=0.9))
    assert O.sizing(1.35, 512, 512, s) == (6, 6)
    assert O.sizing(1.35, 1088, 1088, s) == (14, 14)
    assert O.sizing(1.35, 1024, 1024, s) == (13, 13)
    assert O.sizing(1.35, 1280, 1280, 1.02) == (17, 17)  # clamped sigma
    assert O.sizing(0.8, 2048, 2048, 0.9) == (16, 16)    # test_tcc.sh -> 33 x 33


# ---------------------------------------------------------------- images
PIPE_CASES = [(64, 64, 3, 3), (63, 65, 3, 2), (48, 40, 4, 3), (57, 57, 14, 14), (96, 80, 5, 6)]


@pytest.mark.parametrize("Lx,Ly,Nx,Ny", PIPE_CASES)
def test_socs_closed_form_vs_reference_pipeline(Lx, Ly, Nx, Ny):
    rng = np.random.default_rng(Lx * 1000 + Ly)
    m = (rng.random((Ly, Lx)) > 0.5).astype(np.float32)
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = rng.standard_normal((3, T)).astype(np.float32)
    ki = rng.standard_normal((3, T)).astype(np.float32)
    sc = np.array([1.0, 0.5, 0.25], np.float32)
    a = O.socs_image(m, kr, ki, sc, Nx, Ny)
    b = RP.socs_pipeline(m, kr, ki, sc, Nx, Ny)
    assert relerr(a, b) < 1e-12
    band = O.mask_band(m, Nx, Ny)
    f = RP.ft_r2c(m)
    fb = f[Ly // 2 - Ny:Ly // 2 + Ny + 1, Lx // 2 - Nx:Lx // 2 + Nx + 1]
    assert np.abs(band - fb).max() < 1e-14


def test_fi_rejects_small_period():
    with pytest.raises(ValueError):
        RP.fi(np.zeros((9, 9)), 8, 8)


@pytest.mark.parametrize("L,N", [(32, 2), (31, 2), (40, 3)])
def test_hopkins_vs_reference_pipeline(L, N):
    rng = np.random.default_rng(L)
    T = (2 * N + 1) ** 2
    A = rng.standard_normal((T, T)) + 1j * rng.standard_normal((T, T))
    t = A @ A.conj().T
    m = (rng.random((L, L)) > 0.5).astype(np.float32)
    assert relerr(O.hopkins_image(m, t, N, N), RP.hopkins_pipeline(m, t, N, N)) < 1e-12


# ---------------------------------------------------------------- TCC / eigen
@pytest.fixture(scope="module")
def cfg0():
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(1.35, 512, 512, sig)
    t = O.tcc(src, sig, 1.35, 0.0, 512, 512, Nx, Ny)
    return src, sig, Nx, Ny, t


def test_tcc_closed_forms(cfg0):
    src, sig, Nx, Ny, t = cfg0
    T = t.shape[0]
    assert np.abs(t - t.conj().T).max() == 0.0
    w = np.linalg.eigvalsh(t)
    assert w.min() > -1e-12 * w.max()
    # defocus 0: every used source point passes the pupil at n = 0 -> TCC(0,0) = 1
    c = (T - 1) // 2
    assert abs(t[c, c] - 1.0) < 1e-6
    # trace = sum_s S_s #{n : P(n,s) != 0}
    pup, sv, _ = O.pupil_matrix(src, sig, 1.35, 0.0, 512, 512, Nx, Ny)
    tr = float(np.sum(sv * (np.abs(pup) > 0).sum(1)))
    assert abs(np.trace(t).real - tr) < 1e-10 * tr
    # centrosymmetric source: TCC(-a,-b) == TCC(a,b); ID(-n) = T-1-ID(n)
    assert np.abs(t[::-1, ::-1] - t).max() < 1e-13


def test_point_source_is_coherent():
    src = O.source("Point", 51, px=0.3, py=-0.2)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(0.9, 1024, 1024, sig)
    t = O.tcc(src, sig, 0.9, 0.3, 1024, 1024, Nx, Ny)
    w = np.linalg.eigvalsh(t)
    pup, sv, _ = O.pupil_matrix(src, sig, 0.9, 0.3, 1024, 1024, Nx, Ny)
    assert pup.shape[0] == 1
    npup = int((np.abs(pup) > 0).sum())
    assert npup > 40
    # |P|^2 = cosf^2 + sinf^2 differs from 1 by float rounding only
    assert abs(w[-1] - npup) < 1e-7 * npup
    assert abs(w[-2]) < 1e-9 * npup


@pytest.mark.parametrize("NA,nmed", [(0.9, 1.0), (1.35, 1.44)])
@pytest.mark.parametrize("df", [0.2361, -0.2361, 0.4722, -0.4722])
@pytest.mark.parametrize("pq", [(0, 0), (7, -4), (-16, 9)])
def test_defocus_phase_is_angular_spectrum_propagation(NA, nmed, df, pq):
    """The defocus phase of the pupil (klitho_tcc.cpp:191,197,206-221), pinned by
    physics rather than by the reference's formula: a plane wave of spatial
    frequency nu (cycles/nm) that propagates a distance z in a medium of index n
    gains the phase 2 pi z sqrt(n^2/lambda^2 - |nu|^2) (angular-spectrum
    transfer function).  Here z = defocus * lambda / NA^2 (the normalised
    defocus of the tools' command line), and diffraction order (nx, ny) of a
    period L lit from source cell (p, q) of the sigma-grid (sh cells per unit
    sigma) has nu = (nx, ny)/L + (p, q)/sh * NA/lambda; the pupil passes
    |nu| <= NA/lambda.  With a one-cell source the TCC is P P^H, so
    arg TCC[i][c] = phi(nu_i) - phi(nu_c).  Everything below is float64 in
    physical units; the oracle evaluates the phase in float32 like the
    reference, hence the 1e-5 rad tolerance."""
    lam, n, L = 193.0, 51, 2048
    sh = (n - 1) // 2
    p, q = pq
    src = np.zeros((n, n), np.float32)
    src[q + sh, p + sh] = 1.0
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(NA, L, L, sig)
    t = O.tcc(src, sig, NA, df, L, L, Nx, Ny, nmed=nmed)
    W = 2 * Nx + 1
    ny, nx = np.divmod(np.arange(t.shape[0]), W)
    nx, ny = nx - Nx, ny - Ny
    z = df * lam / NA ** 2
    vx = nx / L + (p / sh) * NA / lam
    vy = ny / L + (q / sh) * NA / lam
    v2 = vx * vx + vy * vy
    kz = np.sqrt(np.maximum(nmed ** 2 / lam ** 2 - v2, 0.0))
    phi = 2.0 * np.pi * z * kz
    rho = np.sqrt(v2) * lam / NA  # |nu| in units of NA / lambda
    c = Ny * W + Nx  # nu_c = the source direction itself, always inside
    col = t[:, c]
    inside = np.abs(col) > 0
    # membership: |nu| <= NA/lambda (decided in float32 by the reference, so
    # samples within 1e-5 of the edge are left out)
    clear = np.abs(rho - 1.0) > 1e-5
    assert np.array_equal(inside[clear], rho[clear] <= 1.0)
    assert inside.sum() > 200
    d = np.angle(col[inside]) - (phi[inside] - phi[c])
    d = (d + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(d).max() < 1e-5
    # the phase varies over the pupil by much more than the tolerance
    assert np.ptp(phi[inside]) > 0.1


def test_relerr_sees_the_imaginary_part():
    """A TCC whose imaginary part has the wrong sign must fail the comparison."""
    rng = np.random.default_rng(0)
    t = rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))
    assert relerr(t.conj(), t) > 0.5
    assert relerr(t, t) == 0.0


def test_pupil_float_decisions_vs_numpy_float32():
    """Membership decided in float32 exactly as klitho_tcc.cpp:206-219."""
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    NA, L = np.float32(0.95), 700
    Nx, Ny = O.sizing(float(NA), L, L, sig)
    pup, sv, pq = O.pupil_matrix(src, sig, float(NA), 0.25, L, L, Nx, Ny)
    f32 = np.float32
    Ln = f32(f32(L) * NA) / f32(193.0)
    sh = 25
    W = 2 * Nx + 1
    for (p, q), row in zip(pq[::37], pup[::37]):
        sx, sy = f32(p) / f32(sh), f32(q) / f32(sh)
        nxmin, nxmax = int((f32(-1) - sx) * Ln), int((f32(1) - sx) * Ln)
        nymin, nymax = int((f32(-1) - sy) * Ln), int((f32(1) - sy) * Ln)
        for ny in range(-Ny, Ny + 1):
            for nx in range(-Nx, Nx + 1):
                fy = f32(f32(ny) / Ln) + sy
                fx = f32(f32(nx) / Ln) + sx
                inside = (nxmin <= nx <= nxmax and nymin <= ny <= nymax
                          and f32(fy * fy) <= 1 and f32(f32(fx * fx) + f32(fy * fy)) <= 1.0)
                assert inside == (abs(row[(ny + Ny) * W + nx + Nx]) > 0)


def test_eigh_vs_lapack_zheevr(cfg0):
    t = cfg0[4]
    w, v = O.eigh_topk(t, 24)
    we, ve = scipy.linalg.eigh(t, driver="evr")
    we, ve = we[::-1], ve[:, ::-1]
    assert np.abs(w - we[:24]).max() < 1e-12 * we[0]
    for i in range(24):  # residual and unit norm
        assert np.linalg.norm(t @ v[i] - w[i] * v[i]) < 1e-11 * we[0]
        assert abs(np.linalg.norm(v[i]) - 1) < 1e-12
    # cluster projectors (annular source gives exact degenerate pairs)
    i = 0
    while i < 24:
        j = i + 1
        while j < 24 and abs(we[j] - we[i]) < 1e-9 * we[0]:
            j += 1
        P1 = v[i:j].T @ v[i:j].conj()
        P2 = ve[:, i:j] @ ve[:, i:j].conj().T
        assert np.abs(P1 - P2).max() < 1e-9
        i = j


def test_eigenvalues_of_independent_restatement(cfg0):
    """The configuration's TCC eigenvalues as an independent FP32-faithful
    restatement of the reference path computed them: lambda_0 = 17.37,
    lambda_23 = 0.0708, lambda_24 = 0.0636, the top 24 carrying 96.0 % of the
    trace."""
    w, _ = O.eigh_topk(cfg0[4], 25)
    assert abs(w[0] - 17.37) < 0.01
    assert abs(w[23] - 0.0708) < 1e-4 and abs(w[24] - 0.0636) < 1e-4
    assert abs(w[:24].sum() / np.trace(cfg0[4]).real - 0.960) < 1e-3


def test_abbe_equals_hopkins_and_socs_all(cfg0):
    src, sig = cfg0[0], cfg0[1]
    L = 256
    Nx, Ny = O.sizing(1.35, L, L, sig)
    t = O.tcc(src, sig, 1.35, 0.0, L, L, Nx, Ny)
    rng = np.random.default_rng(5)
    m = (rng.random((L, L)) > 0.6).astype(np.float32)
    hop = O.hopkins_image(m, t, Nx, Ny)
    abbe = O.abbe_image(m.astype(np.float64), src, 1.35, 0.0, Nx, Ny)
    assert relerr(abbe, hop) < 1e-13
    w, v = O.eigh_topk(t, t.shape[0])
    kr, ki, sc = O.kernels_from_eig(v, w)
    assert relerr(O.socs_image(m, kr, ki, sc, Nx, Ny), hop) < 1e-6  # f32 kernels


# ---------------------------------------------------------------- adjoint
@pytest.fixture(scope="module")
def small_problem():
    rng = np.random.default_rng(11)
    src = O.source("Annular", 21, inner=0.5, outer=0.9)
    sig = O.outer_sigma(src)
    L = 200
    Nx, Ny = O.sizing(1.35, L, L, sig)
    assert (Nx, Ny) == (2, 2)
    t = O.tcc(src, sig, 1.35, 0.0, L, L, Nx, Ny)
    w, v = O.eigh_topk(t, 6)
    kr, ki, sc = O.kernels_from_eig(v, w)
    tgt = (rng.random((L, L)) > 0.5).astype(np.float32)
    m = np.clip(tgt + rng.uniform(-0.1, 0.1, (L, L)), 0, 1)
    return src, Nx, Ny, kr, ki, sc, tgt, m, rng


def test_socs_grad_finite_difference(small_problem):
    src, Nx, Ny, kr, ki, sc, tgt, m, rng = small_problem
    L0, img, g = O.socs_loss_grad(m, tgt, kr, ki, sc, Nx, Ny)
    h = 1e-3  # L ~ 1e4: smaller steps are dominated by rounding in L
    for _ in range(12):
        y, x = rng.integers(0, m.shape[0], 2)
        mp, mm = m.copy(), m.copy()
        mp[y, x] += h
        mm[y, x] -= h
        fd = (O.socs_loss_grad(mp, tgt, kr, ki, sc, Nx, Ny, grad=False)[0]
              - O.socs_loss_grad(mm, tgt, kr, ki, sc, Nx, Ny, grad=False)[0]) / (2 * h)
        assert abs(fd - g[y, x]) < 1e-6 * np.abs(g).max()


def test_socs_grad_homogeneity(small_problem):
    src, Nx, Ny, kr, ki, sc, tgt, m, rng = small_problem
    L0, img, g = O.socs_loss_grad(m, tgt, kr, ki, sc, Nx, Ny)
    s = 1.0 / (1.0 + np.exp(-50.0 * (img - 0.25)))
    G = 2 * (s - tgt) * 50.0 * s * (1 - s)
    assert abs(np.sum(m * g) - 2 * np.sum(G * img)) < 1e-10 * np.sum(np.abs(G * img))
    assert relerr(img, O.socs_image_d(m, kr, ki, sc, Nx, Ny)) < 1e-14


def test_abbe_source_grad_fd_and_identity(small_problem):
    src, Nx, Ny, kr, ki, sc, tgt, m, rng = small_problem
    L0, img, gs = O.abbe_loss_grad(m, tgt, src, 1.35, 0.0, Nx, Ny)
    s = 1.0 / (1.0 + np.exp(-50.0 * (img - 0.25)))
    G = 2 * (s - tgt) * 50.0 * s * (1 - s)
    assert abs(np.sum(src * gs) - np.sum(G * img)) < 1e-9 * np.sum(np.abs(G * img))
    n = src.shape[0]
    pts = [(10, 10), (10, 19), (3, 12), (16, 6), (1, 10)]
    for (q, p) in pts:
        sp, sm = src.copy(), src.copy()
        sp[q, p] = np.float32(src[q, p] + np.float32(1e-4))
        sm[q, p] = np.float32(src[q, p] - np.float32(1e-4))
        d = float(sp[q, p]) - float(sm[q, p])
        fd = (O.abbe_loss_grad(m, tgt, sp, 1.35, 0.0, Nx, Ny, grad=False)[0]
              - O.abbe_loss_grad(m, tgt, sm, 1.35, 0.0, Nx, Ny, grad=False)[0]) / d
        assert abs(fd - gs[q, p]) < 1e-5 * max(1.0, np.abs(gs).max())
    assert gs.shape == (n, n)
