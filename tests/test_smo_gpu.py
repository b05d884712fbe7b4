"""Source-mask optimisation (klt_smo_run) on the GPU.

* Against an oracle-driven loop: the same projected, normalised gradient
  steps with every ingredient from the FP64 oracle (TCC, eigensolve pinned to
  zheevr, SOCS loss and mask gradient, Abbe source gradient; all pinned in
  test_oracle.py): two relaxed 512^2 masks, annular source, K = 24; after each
  of the first three steps the masks and the source agree to relative L2
  1e-4 and the loss to 1e-4.
* Config 4: 16 relaxed 1024^2 Manhattan masks (mask + U(-0.1, 0.1), clipped),
  target = the mask, L = sum (sigmoid(50 (I - 0.25)) - T)^2, K = 24 kernels of
  29 x 29, 51^2 annular source: the loss decreases over 20 steps and the
  printed-resist mismatch count does not grow.
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu

NA, NMED, LAM = 1.35, 1.44, 193.0


def _oracle_loop(m, T, S, steps, L, step_m, step_s, K=24):
    m = m.astype(np.float32).copy()
    S = S.astype(np.float32).copy()
    Nx, Ny = O.sizing(NA, L, L, 1.0)
    n = S.shape[0]
    sh = (n - 1) // 2
    yy, xx = np.mgrid[-sh:n - sh, -sh:n - sh]
    disc = xx * xx + yy * yy <= sh * sh
    out = []
    for it in range(steps + 1):
        t = O.tcc(S, 1.0, NA, 0.0, L, L, Nx, Ny, nmed=NMED)
        w, v = O.eigh_topk(t, K)
        kr, ki, sc = O.kernels_from_eig(v, w)
        tot, grads = 0.0, []
        for b in range(m.shape[0]):
            Lb, _, g = O.socs_loss_grad(m[b].astype(np.float64), T[b], kr, ki, sc, Nx, Ny)
            tot += Lb
            grads.append(g)
        out.append((tot, m.copy(), S.copy()))
        if it == steps:
            break
        gS = np.zeros(S.shape, np.float64)
        for b in range(m.shape[0]):
            gS += O.abbe_loss_grad(m[b].astype(np.float64), T[b], S, NA, 0.0, Nx, Ny,
                                   nmed=NMED)[2]
        for b in range(m.shape[0]):
            g = grads[b]
            m[b] = np.clip(m[b] - step_m * g / np.abs(g).max(), 0.0, 1.0).astype(np.float32)
        Sn = np.where(disc, np.maximum(0.0, S - step_s * S.max() * gS / np.abs(gS).max()), 0.0)
        S = (Sn / Sn.sum()).astype(np.float32)
    return out


def test_smo_matches_oracle_loop(ctx):
    import torch
    print("case smo 2 x 512^2, 3 steps vs oracle loop", flush=True)
    L, steps, sm, ss = 512, 3, 0.05, 0.05
    T = np.stack([W.manhattan(L, 8100 + i) for i in range(2)])
    m0 = np.stack([W.relaxed(T[i], 8200 + i) for i in range(2)])
    S0 = O.source("Annular", 51, inner=0.6, outer=0.9)
    ref = _oracle_loop(m0, T, S0, steps, L, sm, ss)
    dm = torch.from_numpy(m0.copy()).cuda()
    dt = torch.from_numpy(T).cuda()
    S = np.ascontiguousarray(S0.copy())
    for it in range(steps):
        hist, _ = klt.smo_run(ctx, dm, dt, S, NA, 1, step_mask=sm, step_src=ss, nmed=NMED)
        torch.cuda.synchronize()
        assert abs(hist[0] - ref[it][0]) < 1e-4 * abs(ref[it][0]), it
        rl, rm, rS = ref[it + 1]
        assert abs(hist[1] - rl) < 1e-4 * abs(rl), it
        assert relerr(dm.cpu().numpy(), rm) < 1e-4, it
        assert relerr(S, rS) < 1e-4, it
    assert ref[-1][0] < ref[0][0]


def test_smo_config4_loss_decreases(ctx):
    import torch
    print("case smo config 4: 16 x 1024^2, 20 steps", flush=True)
    T = W.manhattan_batch(1024, 16, 3000)
    m0 = np.stack([W.relaxed(T[i], 4000 + i) for i in range(16)])
    dm = torch.from_numpy(m0).cuda()
    dt = torch.from_numpy(T).cuda()
    S = np.ascontiguousarray(O.source("Annular", 51, inner=0.6, outer=0.9))
    hist, mis = klt.smo_run(ctx, dm, dt, S, NA, 20, step_mask=0.05, step_src=0.05, nmed=NMED)
    print("loss", hist.tolist(), "mismatch", mis.tolist(), flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.9 * hist[0]
    assert mis[-1] <= mis[0]
    assert abs(S.sum() - 1.0) < 1e-5 and S.min() >= 0.0
    m = dm.cpu().numpy()
    assert m.min() >= 0.0 and m.max() <= 1.0
