"""Config 4's finite-difference check of the GPU gradients: central finite
differences (step 1e-3) of the FP64 oracle loss on 64 random pixels and 16
source points of one mask, against the GPU's klt_socs_backward (mask) and
klt_abbe_source_grad (source) gradients.

The mask gradient uses config 4's kernels (K = 24 of 29 x 29 from the
51 x 51 annular grid, sigma 0.6 / 0.9, NA 1.35, period 1088 nm) on a relaxed
256 x 256 Manhattan mask instead of 1024 x 1024, so that the 128 FP64 SOCS
loss evaluations of the oracle stay within seconds (the full-size mask
gradient is checked against the oracle's FP64 gradient and by exact identities
in test_adjoint_gpu.py).  The source gradient is checked at config 4's size:
a relaxed 1024 x 1024 mask, the 51 x 51 annular grid, half-size 14.

L = sum_x (sigmoid(50 (I - 0.25)) - T)^2.  Tolerance: |FD - g| <= 1e-4 max|g|
for every sampled entry (central differences of an FP64 loss at step 1e-3 are
accurate to about 1e-7 of max|g| here).
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu
H = 1e-3
TOL = 1e-4


@pytest.fixture(scope="module")
def problem():
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(1.35, W.PERIOD_29, W.PERIOD_29, sig)
    assert (Nx, Ny) == (14, 14)
    t = O.tcc(src, sig, 1.35, 0.0, W.PERIOD_29, W.PERIOD_29, Nx, Ny, nmed=1.44)
    w, v = O.eigh_topk(t, 24)
    kr, ki, sc = O.kernels_from_eig(v, w)
    tgt = W.manhattan(256, 4200)
    m = W.relaxed(tgt, 4201)
    return src, Nx, Ny, kr, ki, sc, tgt, m


def test_mask_gradient_central_differences(ctx, problem):
    print("case config4 FD mask 256 K=24 29x29, 64 pixels", flush=True)
    src, Nx, Ny, kr, ki, sc, tgt, m = problem
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    losses, g, _ = klt.socs_backward(ctx, bank, m[None].copy(), tgt[None].copy())
    bank.close()
    g = np.asarray(g[0], np.float64)
    m64 = m.astype(np.float64)
    rng = np.random.default_rng(4102)
    pix = rng.choice(m.size, 64, replace=False)
    gmax = np.abs(g).max()
    worst = 0.0
    for p in pix:
        y, x = divmod(int(p), m.shape[1])
        mp, mm = m64.copy(), m64.copy()
        mp[y, x] += H
        mm[y, x] -= H
        fd = (O.socs_loss_grad(mp, tgt, kr, ki, sc, Nx, Ny, grad=False)[0]
              - O.socs_loss_grad(mm, tgt, kr, ki, sc, Nx, Ny, grad=False)[0]) / (2 * H)
        worst = max(worst, abs(fd - g[y, x]) / gmax)
    print(f"mask FD: worst |FD - g| / max|g| = {worst:.2e}", flush=True)
    assert worst <= TOL


def test_source_gradient_central_differences(ctx):
    print("case config4 FD source 51x51, 16 points, 1024^2 mask", flush=True)
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    Nx = Ny = 14
    tgt = W.manhattan(1024, 3000)
    m = W.relaxed(tgt, 4000)
    losses, gs, _ = klt.abbe_source_grad(ctx, m[None].copy(), tgt[None].copy(), src, 1.35, 0.0,
                                         Nx, Ny, nmed=1.44)
    gs = np.asarray(gs[0], np.float64)
    _, _, g64 = O.abbe_loss_grad(m.astype(np.float64), tgt, src, 1.35, 0.0, Nx, Ny, nmed=1.44)
    print(f"source gradient rel-L2 vs FP64 {np.linalg.norm(gs - g64) / np.linalg.norm(g64):.2e}, "
          f"max |diff| / max|g| {np.abs(gs - g64).max() / np.abs(g64).max():.2e}", flush=True)
    n = src.shape[0]
    yy, xx = np.mgrid[0:n, 0:n]
    c = (n - 1) / 2
    inside = np.flatnonzero(((yy - c) ** 2 + (xx - c) ** 2 <= c * c).ravel())
    rng = np.random.default_rng(4103)
    pts = rng.choice(inside, 16, replace=False)
    gmax = np.abs(gs).max()
    m64 = m.astype(np.float64)
    worst = 0.0
    for p in pts:
        q, r = divmod(int(p), n)
        sp, sm = src.copy(), src.copy()
        sp[q, r] = np.float32(src[q, r] + np.float32(H))
        sm[q, r] = np.float32(src[q, r] - np.float32(H))
        d = float(sp[q, r]) - float(sm[q, r])
        fd = (O.abbe_loss_grad(m64, tgt, sp, 1.35, 0.0, Nx, Ny, grad=False, nmed=1.44)[0]
              - O.abbe_loss_grad(m64, tgt, sm, 1.35, 0.0, Nx, Ny, grad=False, nmed=1.44)[0]) / d
        worst = max(worst, abs(fd - gs[q, r]) / gmax)
    print(f"source FD: worst |FD - g| / max|g| = {worst:.2e}", flush=True)
    assert worst <= TOL
