"""Source gradient in the Abbe form (klt_abbe_source_grad) for kernel
half-sizes 16..31 (the wide passes: band with 64 twiddle rows, the Hopkins
image formed inside the fused resist band, the 128-point coarse grid), against
the FP64 oracle (whose gradient the CPU suite pins by central finite
differences on the source weights).

The half-size follows from the optics: at a 256^2 period the wavelength is
shortened (40 and 25 nm at NA 1.35) so that the pupil reaches N = 16..31, with
a 21 x 21 annular source grid to keep the oracle's per-point sums short.

Tolerances: source-gradient relative L2 <= 1e-4, loss within 1e-5 (as for
the half-sizes up to 15), and at 1024^2 the identity
sum_s S_s dL/dS_s = sum_x G I_A on the GPU's own outputs.
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu


def _annular(n=21):
    src = O.source("Annular", n, inner=0.6, outer=0.9)
    return src, O.outer_sigma(src)


@pytest.mark.parametrize("lam,df", [(40.0, 0.0), (25.0, 0.2361)])
def test_wide_source_gradient_vs_fp64(ctx, lam, df):
    src, sig = _annular()
    L = 256
    Nx, Ny = O.sizing(1.35, L, L, sig, lam=lam)
    print(f"case abbe wide lam={lam} df={df} N={Nx}", flush=True)
    assert 16 <= Nx <= 31 and Nx == Ny
    tgts = np.stack([W.manhattan(L, 170 + i) for i in range(2)])
    masks = np.stack([W.relaxed(tgts[i], 180 + i) for i in range(2)])
    losses, grads, _ = klt.abbe_source_grad(ctx, masks, tgts, src, 1.35, df, Nx, Ny, nmed=1.44,
                                            lam=lam)
    for i in range(2):
        L0, img, g = O.abbe_loss_grad(masks[i].astype(np.float64), tgts[i], src, 1.35, df, Nx,
                                      Ny, lam=lam, nmed=1.44)
        e = relerr(grads[i], g)
        print(f"mask {i}: gradient rel-L2 {e:.2e}, loss {losses[i]:.6g} vs {L0:.6g}", flush=True)
        assert abs(losses[i] - L0) < 1e-5 * L0
        assert e < 1e-4


def test_wide_source_gradient_identity_1024(ctx):
    import torch
    print("case abbe wide identity 1024", flush=True)
    src, sig = _annular(51)
    Nx, Ny = O.sizing(1.35, 1024, 1024, sig, lam=160.0)
    assert 16 <= Nx <= 31
    t = W.manhattan(1024, 21)
    m = W.relaxed(t, 22)
    dm = torch.from_numpy(m[None].copy()).cuda()
    dt = torch.from_numpy(t[None].copy()).cuda()
    ds = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    losses, grads, images = klt.abbe_source_grad(ctx, dm, dt, ds, 1.35, 0.0, Nx, Ny, nmed=1.44,
                                                 lam=160.0, want_images=True)
    torch.cuda.synchronize()
    ctx.sync()
    I = images[0].double().cpu().numpy()
    g = grads[0].double().cpu().numpy()
    s = 1.0 / (1.0 + np.exp(-50.0 * (I - 0.25)))
    G = 2 * (s - t) * 50.0 * s * (1 - s)
    lhs = float(np.sum(src.astype(np.float64) * g))
    rhs = float(np.sum(G * I))
    print(f"N={Nx}: sum S g {lhs:.8g}, sum G I {rhs:.8g}", flush=True)
    assert abs(lhs - rhs) < 1e-4 * float(np.sum(np.abs(G * I)))
    assert abs(float(losses[0].cpu()) - float(np.sum((s - t) ** 2))) < 1e-4 * float(losses[0].cpu())
