"""The vector-Jacobian product of the SOCS image (klt_socs_vjp): dL/dm for a
given dL/dI, kernel half-sizes up to 31.

The image is quadratic in the mask, so central differences of sum_x G(x) I(x)
are exact up to rounding: the FP64 oracle image differenced at step 1e-3 pins
dL/dm for any G.  The checks:
  * random G (normal, and scaled by 2^-40 and 2^40: the per-tile scaling of
    the split keeps both halves in range, so the result scales exactly):
    central differences of the oracle at 32 pixels, |FD - g| <= 1e-4 max|g|;
  * G = dL/dI of the resist loss, computed in FP64 from the oracle image: the
    product equals the oracle's resist-loss gradient to relative L2 1e-4,
    N = 14 (even and odd sizes) and N = 20 (the wide passes);
  * at 1024^2, <m, dL/dm> = 2 sum_x G I (I is homogeneous of degree 2 in m),
    from the GPU's own images.
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu
H = 1e-3


def _bank(rng, Nx, Ny, nk):
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, nk))[::-1].astype(np.float32).copy()
    return kr, ki, sc


@pytest.mark.parametrize("N", [14, 20])
def test_vjp_central_differences(ctx, N):
    print(f"case vjp FD N={N}", flush=True)
    rng = np.random.default_rng(300 + N)
    kr, ki, sc = _bank(rng, N, N, 5)
    bank = ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1)
    L = 256
    m = W.relaxed(W.manhattan(L, 310 + N), 311 + N)
    G = rng.standard_normal((L, L)).astype(np.float32)
    g = np.asarray(klt.socs_vjp(ctx, bank, m[None].copy(), G[None].copy())[0], np.float64)
    for c in (2.0 ** -40, 2.0 ** 40):
        gc = np.asarray(klt.socs_vjp(ctx, bank, m[None].copy(), (G * c)[None].copy())[0],
                        np.float64)
        assert relerr(gc / c, g) < 1e-6, c
    m64 = m.astype(np.float64)
    G64 = G.astype(np.float64)
    gmax = np.abs(g).max()
    worst = 0.0
    for p in rng.choice(L * L, 32, replace=False):
        y, x = divmod(int(p), L)
        mp, mm = m64.copy(), m64.copy()
        mp[y, x] += H
        mm[y, x] -= H
        fd = (np.sum(G64 * O.socs_image_d(mp, kr, ki, sc, N, N))
              - np.sum(G64 * O.socs_image_d(mm, kr, ki, sc, N, N))) / (2 * H)
        worst = max(worst, abs(fd - g[y, x]) / gmax)
    print(f"N={N}: worst |FD - g| / max|g| = {worst:.2e}", flush=True)
    assert worst <= 1e-4
    bank.close()


@pytest.mark.parametrize("Lx,Ly,N", [(256, 256, 14), (260, 255, 14), (300, 257, 20)])
def test_vjp_resist_gradient(ctx, Lx, Ly, N):
    print(f"case vjp resist {Lx}x{Ly} N={N}", flush=True)
    rng = np.random.default_rng(Lx + Ly + N)
    kr, ki, sc = _bank(rng, N, N, 6)
    bank = ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1)
    t = (rng.random((2, Ly, Lx)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    refs, Gs = [], []
    for i in range(2):
        img = O.socs_image_d(m[i].astype(np.float64), kr, ki, sc, N, N)
        a, thr = 4.0 / float(np.std(img)), float(np.median(img))
        s = 1.0 / (1.0 + np.exp(-a * (img - thr)))
        Gs.append((2 * (s - t[i]) * a * s * (1 - s)).astype(np.float32))
        refs.append(O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, N, N, a=a,
                                     thr=thr)[2])
    g = klt.socs_vjp(ctx, bank, m, np.stack(Gs))
    for i in range(2):
        e = relerr(g[i], refs[i])
        print(f"mask {i}: rel-L2 vs the oracle's resist-loss gradient {e:.2e}", flush=True)
        assert e < 1e-4
    bank.close()


def test_vjp_identity_1024(ctx):
    import torch
    print("case vjp identity 1024", flush=True)
    rng = np.random.default_rng(1024)
    kr, ki, sc = _bank(rng, 14, 14, 24)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    m = np.stack([W.relaxed(W.manhattan(1024, 520 + i), 530 + i) for i in range(2)])
    G = rng.standard_normal(m.shape).astype(np.float32)
    dm, dG = torch.from_numpy(m).cuda(), torch.from_numpy(G).cuda()
    img = ctx.socs_forward(bank, dm)
    g = klt.socs_vjp(ctx, bank, dm, dG)
    ctx.sync()
    I = img.double().cpu().numpy()
    gg = g.double().cpu().numpy()
    for i in range(2):
        lhs, rhs = float(np.sum(m[i] * gg[i])), float(2 * np.sum(G[i] * I[i]))
        print(f"mask {i}: <m, g> {lhs:.8g}, 2 sum G I {rhs:.8g}", flush=True)
        assert abs(lhs - rhs) < 1e-4 * float(np.sum(np.abs(G[i] * I[i])))
    bank.close()
