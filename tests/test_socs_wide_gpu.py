"""The batched matrix-core SOCS forward for kernel half-sizes 16..31
(klt_socs_forward: band pass with 64 twiddle rows, coarse stage on the
128-point grid, synthesis at K = 128 in k_synth_wide).

Tolerance: image relative L2 <= 1e-5 against the FP64 oracle (split-fp16
products with FP32 accumulation), as for the half-sizes up to 15.  The
checks:
  * the route: a call at N = 16 runs k_band_mfma and k_synth_wide, not the
    plain path's passes;
  * the reference's own shipped 33 x 33 kernels (N = 16, three kernels) on
    two 1024^2 Manhattan masks, each image against the oracle;
  * N = 31 (the largest half-size) on unequal, odd sizes, and Nx != Ny;
  * the smallest accepted period (4N+1) for the forward, the adjoint and the
    product with a given dL/dI; 600 masks in sub-batches of 512;
  * every pinned tile height / split factor gives the same image to 1e-6
    (the splits only reorder FP32 sums; the row ranges write disjoint rows),
    and the first matches the oracle.
"""
import itertools
import os

import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-5
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _random_bank(rng, Nx, Ny, nk):
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, nk))[::-1].astype(np.float32).copy()
    return kr, ki, sc


def _golden_bank():
    sx, sy, sc = klt.read_kernel_info(os.path.join(GOLD, "kernel_info.txt"))
    assert (sx, sy, len(sc)) == (33, 33, 3)
    kr = np.stack([np.fromfile(os.path.join(GOLD, f"krn_{i}_r.bin"), np.float32)
                   for i in range(3)])
    ki = np.stack([np.fromfile(os.path.join(GOLD, f"krn_{i}_i.bin"), np.float32)
                   for i in range(3)])
    return kr, ki, np.asarray(sc, np.float32)


def test_wide_route(ctx):
    import torch
    print("case wide route N=16 256^2", flush=True)
    rng = np.random.default_rng(16)
    kr, ki, sc = _random_bank(rng, 16, 16, 4)
    bank = ctx.kernels(kr, ki, sc, 33, 33)
    m = torch.from_numpy((rng.random((2, 256, 256)) > 0.5).astype(np.float32)).cuda()
    ctx.set_profiling(True)
    try:
        img = ctx.socs_forward(bank, m)
        ctx.sync()
        names = set(ctx.profile_report())
    finally:
        ctx.set_profiling(False)
    print("kernels:", sorted(names), flush=True)
    assert {"k_band_mfma", "k_band_assemble", "k_synth_wide"} <= names
    assert not names & {"k_band_cols", "k_band_rows", "k_synth_cols", "k_synth_rows"}
    img = img.cpu().numpy()
    mm = m.cpu().numpy()
    for i in range(2):
        assert relerr(img[i], O.socs_image(mm[i], kr, ki, sc, 16, 16)) < TOL
    bank.close()


def test_reference_kernels_1024(ctx):
    import torch
    print("case wide reference 33x33 kernels, two 1024^2 masks", flush=True)
    kr, ki, sc = _golden_bank()
    bank = ctx.kernels(kr, ki, sc, 33, 33)
    m = np.stack([W.manhattan(1024, 5100), W.relaxed(W.manhattan(1024, 5101), 5102)])
    img = ctx.socs_forward(bank, torch.from_numpy(m).cuda())
    ctx.sync()
    img = img.cpu().numpy()
    for i in range(2):
        e = relerr(img[i], O.socs_image(m[i], kr, ki, sc, 16, 16))
        print(f"image {i}: rel-L2 vs FP64 {e:.2e}", flush=True)
        assert e < TOL
    bank.close()


@pytest.mark.parametrize("Lx,Ly,Nx,Ny", [(1024, 768, 31, 31), (385, 257, 31, 29),
                                         (512, 300, 16, 9), (200, 640, 7, 27)])
def test_wide_sizes(ctx, Lx, Ly, Nx, Ny):
    print(f"case wide Lx={Lx} Ly={Ly} Nx={Nx} Ny={Ny}", flush=True)
    rng = np.random.default_rng(Lx + 3 * Ly + Nx)
    kr, ki, sc = _random_bank(rng, Nx, Ny, 6)
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    m = (rng.random((2, Ly, Lx)) > 0.55).astype(np.float32)
    img = ctx.socs_forward(bank, m)
    for i in range(2):
        e = relerr(img[i], O.socs_image(m[i], kr, ki, sc, Nx, Ny))
        print(f"image {i}: rel-L2 vs FP64 {e:.2e}", flush=True)
        assert e < TOL
    bank.close()


def test_wide_pinned_splits(ctx):
    import torch
    rng = np.random.default_rng(77)
    kr, ki, sc = _random_bank(rng, 20, 20, 5)
    bank = ctx.kernels(kr, ki, sc, 41, 41)
    m = (rng.random((2, 512, 512)) > 0.5).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    base = None
    try:
        for rb, ks, rs in itertools.product((128, 64), (1, 4), (1, 3, 16)):
            print(f"case wide pinned rows={rb} ks={ks} rs={rs}", flush=True)
            ctx.set_splits(ks, rs, 0)
            ctx.set_band_rows(rb)
            out = ctx.socs_forward(bank, dm)
            ctx.sync()
            img = out.cpu().numpy()
            if base is None:
                base = img
                for i in range(2):
                    assert relerr(img[i], O.socs_image(m[i], kr, ki, sc, 20, 20)) < TOL
                continue
            for i in range(2):
                assert relerr(img[i], base[i]) < 1e-6, (rb, ks, rs, i)
    finally:
        ctx.set_splits(0, 0, 0)
        ctx.set_band_rows(0)
    bank.close()


@pytest.mark.parametrize("N", [16, 31])
def test_wide_minimum_size(ctx, N):
    """The smallest period the reference accepts, 4N+1 (klitho_socs.cpp:164-169):
    one partly filled 128-column strip and band row block; forward, adjoint and
    the product with a given dL/dI against the oracle."""
    L = 4 * N + 1
    print(f"case wide minimum size N={N} L={L}", flush=True)
    rng = np.random.default_rng(L)
    kr, ki, sc = _random_bank(rng, N, N, 4)
    bank = ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1)
    t = (rng.random((2, L, L)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    img = ctx.socs_forward(bank, m)
    a, thr = 4.0 / float(np.std(img[0])), float(np.median(img[0]))
    lo, g, _ = klt.socs_backward(ctx, bank, m, t, a=a, thr=thr)
    G = rng.standard_normal(m.shape).astype(np.float32)
    gv = klt.socs_vjp(ctx, bank, m, G)
    for i in range(2):
        assert relerr(img[i], O.socs_image(m[i], kr, ki, sc, N, N)) < TOL
        Lr, _, gr = O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, N, N, a=a,
                                     thr=thr)
        assert relerr(g[i], gr) < 1e-4
        assert abs(lo[i] - Lr) < 1e-5 * Lr
        # <m, dL/dm> = 2 sum G I for the product with G (I quadratic in m)
        I = O.socs_image_d(m[i].astype(np.float64), kr, ki, sc, N, N)
        lhs, rhs = float(np.sum(m[i] * gv[i])), float(2 * np.sum(G[i] * I))
        assert abs(lhs - rhs) < 1e-4 * float(np.sum(np.abs(G[i] * I)))
    bank.close()


def test_wide_sub_batches(ctx):
    """600 masks of 72 x 80 at N = 16: the forward and the adjoint run in
    sub-batches of 512 that share the context's scratch; images and gradients
    on both sides of the boundary against the oracle."""
    print("case wide sub-batches 600 masks", flush=True)
    rng = np.random.default_rng(600)
    kr, ki, sc = _random_bank(rng, 16, 16, 3)
    bank = ctx.kernels(kr, ki, sc, 33, 33)
    t = (rng.random((600, 72, 80)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    img = ctx.socs_forward(bank, m)
    a, thr = 4.0 / float(np.std(img[0])), float(np.median(img[0]))
    lo, g, _ = klt.socs_backward(ctx, bank, m, t, a=a, thr=thr)
    for i in (0, 511, 512, 599):
        assert relerr(img[i], O.socs_image(m[i], kr, ki, sc, 16, 16)) < TOL, i
        Lr, _, gr = O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, 16, 16, a=a,
                                     thr=thr)
        assert relerr(g[i], gr) < 1e-4, i
        assert abs(lo[i] - Lr) < 1e-5 * Lr, i
    bank.close()
