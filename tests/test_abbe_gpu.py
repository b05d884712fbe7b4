"""Source gradient in the Abbe form (klt_abbe_source_grad) against the FP64
oracle, whose gradient the CPU suite pins by central finite differences on
the source weights and by the identity sum_s S_s dL/dS_s = sum_x G I_A.

Tolerances: source-gradient relative L2 <= 1e-4, loss within 1e-5, and the
same identity on the GPU's own outputs at the full 1024 x 1024 size.
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


@pytest.mark.parametrize("L,df,N", [(256, 0.0, None), (512, 0.2361, None), (300, 0.0, 6),
                                    (255, 0.2361, 6)])
def test_source_gradient_vs_fp64(ctx, L, df, N):
    print(f"case abbe L={L} df={df}", flush=True)
    src, sig = _annular()
    Nx, Ny = O.sizing(1.35, L, L, sig) if N is None else (N, N)
    tgts = np.stack([W.manhattan(L, 70 + i) for i in range(2)])
    masks = np.stack([W.relaxed(tgts[i], 80 + i) for i in range(2)])
    losses, grads, _ = klt.abbe_source_grad(ctx, masks, tgts, src, 1.35, df, Nx, Ny, nmed=1.44)
    for i in range(2):
        L0, img, g = O.abbe_loss_grad(masks[i].astype(np.float64), tgts[i], src, 1.35, df, Nx,
                                      Ny, nmed=1.44)
        assert abs(losses[i] - L0) < 1e-5 * L0
        assert relerr(grads[i], g) < 1e-4, relerr(grads[i], g)


def test_source_gradient_full_size_identity(ctx):
    import torch
    print("case abbe identity 1024", flush=True)
    src, sig = _annular()
    t = W.manhattan(1024, 11)
    m = W.relaxed(t, 12)
    dm = torch.from_numpy(m[None].copy()).cuda()
    dt = torch.from_numpy(t[None].copy()).cuda()
    ds = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    losses, grads, images = klt.abbe_source_grad(ctx, dm, dt, ds, 1.35, 0.0, 14, 14, nmed=1.44,
                                                 want_images=True)
    torch.cuda.synchronize()
    ctx.sync()
    I = images[0].double().cpu().numpy()
    g = grads[0].double().cpu().numpy()
    s = 1.0 / (1.0 + np.exp(-50.0 * (I - 0.25)))
    G = 2 * (s - t) * 50.0 * s * (1 - s)
    lhs = float(np.sum(src.astype(np.float64) * g))
    rhs = float(np.sum(G * I))
    assert abs(lhs - rhs) < 1e-4 * float(np.sum(np.abs(G * I)))
    # points outside the unit disc carry no gradient; points with S = 0 inside do
    yy, xx = np.mgrid[-25:26, -25:26]
    assert np.all(g[xx * xx + yy * yy > 625] == 0)
    assert np.abs(g[(xx * xx + yy * yy <= 625) & (src == 0)]).max() > 0
