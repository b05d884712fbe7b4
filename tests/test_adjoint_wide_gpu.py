"""Adjoint of the SOCS image under the resist loss for kernel half-sizes
16..31 (klt_socs_backward on the wide passes: fused synthesis + resist band
at K = 128, 128-point coarse stages).

L = sum_x (sigmoid(a (I - thr)) - T)^2.  The GPU results are compared with
the FP64 oracle's (whose gradient the CPU suite pins by central finite
differences): images to relative L2 1e-5, losses to 1e-5, gradients to
relative L2 1e-4 (the bars of the half-sizes up to 15).  For random kernels
thr is the median of the oracle image and a = 4 / its standard deviation, so
that the sigmoid is neither flat nor saturated.  The checks:
  * the reference's shipped 33 x 33 kernels (N = 16) at 256^2;
  * random kernels at N = 20, N = 31 and Nx != Ny, odd and unequal sizes;
  * pinned tile heights / split factors / kernel groups agree to 1e-5 and
    match the oracle;
  * at 1024^2 the exact identity <m, dL/dm> = 2 sum_x G I (I is quadratic in
    m), G = dL/dI, from the GPU's own outputs.
"""
import os

import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu
GTOL = 1e-4
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _golden():
    sx, sy, sc = klt.read_kernel_info(os.path.join(GOLD, "kernel_info.txt"))
    kr = np.stack([np.fromfile(os.path.join(GOLD, f"krn_{i}_r.bin"), np.float32)
                   for i in range(3)])
    ki = np.stack([np.fromfile(os.path.join(GOLD, f"krn_{i}_i.bin"), np.float32)
                   for i in range(3)])
    return kr, ki, np.asarray(sc, np.float32), 16, 16


def _random(rng, Nx, Ny, nk):
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, nk))[::-1].astype(np.float32).copy()
    return kr, ki, sc, Nx, Ny


def _masks(rng, B, Ly, Lx):
    t = (rng.random((B, Ly, Lx)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    return m, t


def _sigmoid_params(m, kr, ki, sc, Nx, Ny):
    img = O.socs_image(m, kr, ki, sc, Nx, Ny)
    return 4.0 / float(np.std(img)), float(np.median(img))


CASES = [("golden", 256, 256), ("r20", 256, 256), ("r31", 300, 257), ("r16x9", 260, 255)]


@pytest.mark.parametrize("kind,Lx,Ly", CASES)
def test_wide_gradient_vs_fp64(ctx, kind, Lx, Ly):
    print(f"case wide adjoint {kind} {Lx}x{Ly}", flush=True)
    rng = np.random.default_rng(Lx * 3 + Ly)
    if kind == "golden":
        kr, ki, sc, Nx, Ny = _golden()
    elif kind == "r20":
        kr, ki, sc, Nx, Ny = _random(rng, 20, 20, 5)
    elif kind == "r31":
        kr, ki, sc, Nx, Ny = _random(rng, 31, 31, 4)
    else:
        kr, ki, sc, Nx, Ny = _random(rng, 16, 9, 6)
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    m, t = _masks(rng, 2, Ly, Lx)
    a, thr = _sigmoid_params(m[0], kr, ki, sc, Nx, Ny)
    losses, grads, images = klt.socs_backward(ctx, bank, m, t, a=a, thr=thr, want_images=True)
    for i in range(2):
        L0, img, g = O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, Nx, Ny, a=a,
                                      thr=thr)
        ei, eg = relerr(images[i], img), relerr(grads[i], g)
        print(f"image {i}: image {ei:.2e}, gradient {eg:.2e}, loss {losses[i]:.6g} vs {L0:.6g}",
              flush=True)
        assert ei < 1e-5
        assert abs(losses[i] - L0) < 1e-5 * L0
        assert eg < GTOL
    bank.close()


def test_wide_backward_pinned_splits(ctx):
    import torch
    rng = np.random.default_rng(2020)
    kr, ki, sc, Nx, Ny = _random(rng, 20, 20, 6)
    bank = ctx.kernels(kr, ki, sc, 41, 41)
    m, t = _masks(rng, 2, 512, 512)
    a, thr = _sigmoid_params(m[0], kr, ki, sc, Nx, Ny)
    dm, dt = torch.from_numpy(m).cuda(), torch.from_numpy(t).cuda()
    ref = [O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, Nx, Ny, a=a, thr=thr)
           for i in range(2)]
    base = None
    try:
        for ks, rs, kg, rb in ((1, 1, 1, 128), (4, 3, 2, 64), (2, 16, 6, 128), (8, 6, 3, 64)):
            print(f"case wide backward ks={ks} rs={rs} kg={kg} rows={rb}", flush=True)
            ctx.set_splits(ks, rs, kg)
            ctx.set_band_rows(rb)
            lo, g, _ = klt.socs_backward(ctx, bank, dm, dt, a=a, thr=thr)
            ctx.sync()
            g, lo = g.cpu().numpy(), lo.cpu().numpy()
            for i in range(2):
                Lr, _, gr = ref[i]
                assert relerr(g[i], gr) < GTOL, (ks, rs, kg, rb, i)
                assert abs(lo[i] - Lr) < 1e-5 * Lr, (ks, rs, kg, rb, i)
                if base is not None:
                    assert relerr(g[i], base[i]) < 1e-5, (ks, rs, kg, rb, i)
            if base is None:
                base = g
    finally:
        ctx.set_splits(0, 0, 0)
        ctx.set_band_rows(0)
    bank.close()


def test_wide_identity_1024(ctx):
    import torch
    print("case wide adjoint identity 1024, reference kernels", flush=True)
    kr, ki, sc, Nx, Ny = _golden()
    bank = ctx.kernels(kr, ki, sc, 33, 33)
    t = np.stack([W.manhattan(1024, 610), W.manhattan(1024, 611)])
    m = np.stack([W.relaxed(t[0], 612), W.relaxed(t[1], 613)])
    a, thr = _sigmoid_params(m[0], kr, ki, sc, Nx, Ny)
    dm, dt = torch.from_numpy(m).cuda(), torch.from_numpy(t).cuda()
    losses, grads, images = klt.socs_backward(ctx, bank, dm, dt, a=a, thr=thr, want_images=True)
    ctx.sync()
    for i in range(2):
        I = images[i].double().cpu().numpy()
        g = grads[i].double().cpu().numpy()
        s = 1.0 / (1.0 + np.exp(-a * (I - thr)))
        G = 2 * (s - t[i]) * a * s * (1 - s)
        lhs, rhs = float(np.sum(m[i] * g)), float(2 * np.sum(G * I))
        print(f"image {i}: <m, g> {lhs:.8g}, 2 sum G I {rhs:.8g}", flush=True)
        assert abs(lhs - rhs) < 1e-4 * np.sum(np.abs(G * I))
        Lgpu = float(losses[i].cpu())
        assert abs(Lgpu - float(np.sum((s - t[i]) ** 2))) < 1e-4 * Lgpu
    bank.close()
