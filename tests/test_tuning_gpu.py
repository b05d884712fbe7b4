"""Split factors of the batched SOCS passes (klt_ctx_set_autotune).

Small batches run the band pass over several column ranges per row block,
the synthesis pass over several row ranges per strip and the coarse-grid
stage over several groups of kernels. Every split must give the same images
and gradients: the column and kernel-group partials only change the order of
FP32 sums, and the row ranges write disjoint rows. The tests check:
  * forward: images with the fixed-rule splits (tuning off) and with the tuned
    splits agree to relative L2 1e-6, and both match the FP64 oracle to 1e-5;
  * backward: the same for gradients (1e-6 between splits, 1e-4 vs FP64) and
    losses;
  * the tuned shapes are listed by klt_ctx_tune_report.
"""
import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu


def _bank(ctx, n=14, nk=6, seed=5):
    rng = np.random.default_rng(seed)
    T = (2 * n + 1) ** 2
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, nk))[::-1].astype(np.float32).copy()
    return ctx.kernels(kr, ki, sc, 2 * n + 1, 2 * n + 1), (kr, ki, sc)


@pytest.mark.parametrize("L", [256, 384])
def test_forward_splits_agree(ctx, L):
    import torch
    print(f"case tuning forward L={L}", flush=True)
    bank, (kr, ki, sc) = _bank(ctx)
    rng = np.random.default_rng(L)
    m = (rng.random((2, L, L)) > 0.6).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    # the library works on its own stream: wait for it before reading results
    ctx.set_autotune(False)
    out = ctx.socs_forward(bank, dm)
    ctx.sync()
    fixed = out.cpu().numpy()
    ctx.set_autotune(True)
    out = ctx.socs_forward(bank, dm)  # tunes this shape
    ctx.sync()
    tuned = out.cpu().numpy()
    out = ctx.socs_forward(bank, dm)  # uses the tuned splits
    ctx.sync()
    again = out.cpu().numpy()
    for i in range(2):
        assert relerr(tuned[i], fixed[i]) < 1e-6
        assert relerr(again[i], tuned[i]) < 1e-6
        assert relerr(tuned[i], O.socs_image(m[i], kr, ki, sc, 14, 14)) < 1e-5
    shapes = [r for r in ctx.tune_report()
              if r[0] == "socs_forward" and r[1:6] == (2, L, L, 14, 14)]
    assert len(shapes) == 1 and shapes[0][7] in (1, 2, 4, 8)
    assert shapes[0][8] in (1, 2, 3, 4, 6, 8, 12, 16)  # synthesis row ranges
    assert shapes[0][9] in (1, 2, 4)  # kernel groups (candidates up to the 6 kernels)
    assert shapes[0][10] in (64, 128)  # rows per band block


def test_backward_splits_agree(ctx):
    import klt
    import torch
    print("case tuning backward", flush=True)
    L = 256
    bank, (kr, ki, sc) = _bank(ctx, seed=9)
    rng = np.random.default_rng(21)
    t = (rng.random((2, L, L)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    dt = torch.from_numpy(t).cuda()
    ctx.set_autotune(False)
    l0, g0, _ = klt.socs_backward(ctx, bank, dm, dt, thr=0.02)
    ctx.sync()
    ctx.set_autotune(True)
    l1, g1, _ = klt.socs_backward(ctx, bank, dm, dt, thr=0.02)
    ctx.sync()
    g0, g1 = g0.cpu().numpy(), g1.cpu().numpy()
    l0, l1 = l0.cpu().numpy(), l1.cpu().numpy()
    for i in range(2):
        assert relerr(g1[i], g0[i]) < 1e-6
        assert abs(l1[i] - l0[i]) < 1e-6 * abs(l0[i])
        Lr, _, gr = O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, 14, 14, thr=0.02)
        assert relerr(g1[i], gr) < 1e-4
        assert abs(l1[i] - Lr) < 1e-5 * Lr
    assert any(r[0] == "socs_backward" and r[1:6] == (2, L, L, 14, 14) for r in ctx.tune_report())
