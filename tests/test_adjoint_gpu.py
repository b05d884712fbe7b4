"""Adjoint of the SOCS image under the resist loss (klt_socs_backward).

L = sum_x (sigmoid(50 (I - 0.25)) - T)^2 on relaxed Manhattan masks
(mask + U(-0.1, 0.1), clipped to [0, 1]; T = the binary mask).  The GPU
gradient is compared with the FP64 oracle's pixel-by-pixel gradient (which
the CPU suite pins by central finite differences, step 1e-3):
  * gradient relative L2 <= 1e-4 (the adjoint target), loss within 1e-5;
  * odd and unequal periods (mask-spectrum centre floor(L/2) vs image
    centre (L+1)/2);
  * at the full 1024 x 1024 size, the exact identity <m, dL/dm> = 2 sum_x G I
    (I is quadratic in m), G = dL/dI, from the GPU's own outputs.
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu
GTOL = 1e-4


def _bank(period=W.PERIOD_29, nk=24):
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(1.35, period, period, sig)
    t = O.tcc(src, sig, 1.35, 0.0, period, period, Nx, Ny, nmed=1.44)
    w, v = O.eigh_topk(t, nk)
    kr, ki, sc = O.kernels_from_eig(v, w)
    return kr, ki, sc, Nx, Ny


@pytest.fixture(scope="module")
def bank29():
    return _bank()


def _problem(L, seed):
    tgt = W.manhattan(L, seed)
    return W.relaxed(tgt, seed + 7), tgt


@pytest.mark.parametrize("L", [256, 512])
def test_mask_gradient_vs_fp64(ctx, bank29, L):
    print(f"case adjoint L={L}", flush=True)
    kr, ki, sc, Nx, Ny = bank29
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    masks, tgts = zip(*[_problem(L, 900 + i) for i in range(2)])
    masks = np.ascontiguousarray(np.stack(masks))
    tgts = np.ascontiguousarray(np.stack(tgts))
    losses, grads, images = klt.socs_backward(ctx, bank, masks, tgts, want_images=True)
    for i in range(2):
        L0, img, g = O.socs_loss_grad(masks[i].astype(np.float64), tgts[i], kr, ki, sc, Nx, Ny)
        assert relerr(images[i], img) < 1e-5
        assert abs(losses[i] - L0) < 1e-5 * L0
        assert relerr(grads[i], g) < GTOL, relerr(grads[i], g)


@pytest.mark.parametrize("Lx,Ly", [(260, 255), (255, 257)])
def test_mask_gradient_odd_periods(ctx, Lx, Ly):
    print(f"case adjoint odd {Lx}x{Ly}", flush=True)
    rng = np.random.default_rng(Lx + Ly)
    Nx, Ny = 6, 5
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = (rng.standard_normal((4, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((4, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.array([8.0, 4.0, 2.0, 1.0], np.float32)
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    tgt = (rng.random((Ly, Lx)) > 0.6).astype(np.float32)
    m = np.clip(tgt + rng.uniform(-0.1, 0.1, tgt.shape), 0, 1).astype(np.float32)
    losses, grads, _ = klt.socs_backward(ctx, bank, m[None], tgt[None], thr=0.02)
    L0, img, g = O.socs_loss_grad(m.astype(np.float64), tgt, kr, ki, sc, Nx, Ny, thr=0.02)
    assert abs(losses[0] - L0) < 1e-5 * L0
    assert relerr(grads[0], g) < GTOL


def test_homogeneity_full_size_device(ctx, bank29):
    import torch
    print("case adjoint identity 1024", flush=True)
    kr, ki, sc, Nx, Ny = bank29
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    m, t = _problem(1024, 1234)
    dm = torch.from_numpy(m[None].copy()).cuda()
    dt = torch.from_numpy(t[None].copy()).cuda()
    losses, grads, images = klt.socs_backward(ctx, bank, dm, dt, want_images=True)
    torch.cuda.synchronize()
    ctx.sync()
    I = images[0].double().cpu().numpy()
    g = grads[0].double().cpu().numpy()
    s = 1.0 / (1.0 + np.exp(-50.0 * (I - 0.25)))
    G = 2 * (s - t) * 50.0 * s * (1 - s)
    lhs, rhs = float(np.sum(m * g)), float(2 * np.sum(G * I))
    assert abs(lhs - rhs) < 1e-4 * np.sum(np.abs(G * I))
    assert abs(float(losses[0].cpu()) - float(np.sum((s - t) ** 2))) < 1e-4 * float(losses[0].cpu())
