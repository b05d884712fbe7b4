"""Pinned split factors (klt_ctx_set_splits) and band tile heights
(klt_ctx_set_band_rows) of the batched SOCS passes.

The tuner only keeps the choice that runs fastest, so the other candidates'
results are never read back. These tests pin every factor explicitly at
1024 x 1024 (8 or 16 band row blocks, 8 column chunks, 32 synthesis row
groups):
  * forward, 2 masks, 24 kernels: band column ranges ks in {1, 2, 4, 8},
    synthesis row ranges rs in {1, 4, 6, 16} (6 leaves a shorter last range),
    kernel groups kg in {1, 5, 24} (5 leaves groups of unequal size), band
    blocks of 128 and 64 rows. Every
    combination gives the same images as ks = rs = kg = 1 to relative L2 1e-6
    (the splits only reorder FP32 sums), and that image matches the FP64
    oracle to 1e-5;
  * backward, 2 relaxed masks: (ks, rs, kg, rows) in {(1, 1, 1, 128),
    (8, 16, 24, 64), (2, 4, 5, 64), (2, 3, 5, 128), (4, 6, 12, 128)}:
    gradients agree to 1e-5 and match the FP64 oracle to 1e-4, losses to
    1e-5;
  * 64-row band blocks at odd, unequal sizes (the last block partly past the
    image), forward and backward, against the oracle;
  * factors beyond a shape's limits are lowered, not rejected: ks = 64 at
    256 x 256 (2 column chunks) runs and equals ks = 2.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu
N, NK, L = 14, 24, 1024


def _bank(ctx, seed):
    rng = np.random.default_rng(seed)
    T = (2 * N + 1) ** 2
    kr = (rng.standard_normal((NK, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((NK, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, NK))[::-1].astype(np.float32).copy()
    return ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1), (kr, ki, sc)


@pytest.fixture
def pinned(ctx):
    yield ctx
    ctx.set_splits(0, 0, 0)
    ctx.set_band_rows(0)


def test_forward_pinned_splits(pinned):
    import torch
    ctx = pinned
    bank, (kr, ki, sc) = _bank(ctx, 31)
    rng = np.random.default_rng(8)
    m = (rng.random((2, L, L)) > 0.55).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    base = None
    for rb, ks, rs, kg in itertools.product((128, 64), (1, 2, 4, 8), (1, 4, 6, 16), (1, 5, 24)):
        print(f"case forward rows={rb} ks={ks} rs={rs} kg={kg}", flush=True)
        ctx.set_splits(ks, rs, kg)
        ctx.set_band_rows(rb)
        out = ctx.socs_forward(bank, dm)
        ctx.sync()
        img = out.cpu().numpy()
        if base is None:
            base = img
            for i in range(2):
                assert relerr(img[i], O.socs_image(m[i], kr, ki, sc, N, N)) < 1e-5
            continue
        for i in range(2):
            assert relerr(img[i], base[i]) < 1e-6, (rb, ks, rs, kg, i)
    bank.close()


def test_backward_pinned_splits(pinned):
    import klt
    import torch
    ctx = pinned
    bank, (kr, ki, sc) = _bank(ctx, 37)
    rng = np.random.default_rng(9)
    t = (rng.random((2, L, L)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    dm = torch.from_numpy(m).cuda()
    dt = torch.from_numpy(t).cuda()
    ref = [O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, N, N, thr=0.02)
           for i in range(2)]
    base = None
    for ks, rs, kg, rb in ((1, 1, 1, 128), (8, 16, 24, 64), (2, 4, 5, 64), (2, 3, 5, 128),
                           (4, 6, 12, 128)):
        print(f"case backward ks={ks} rs={rs} kg={kg} rows={rb}", flush=True)
        ctx.set_splits(ks, rs, kg)
        ctx.set_band_rows(rb)
        lo, g, _ = klt.socs_backward(ctx, bank, dm, dt, thr=0.02)
        ctx.sync()
        g, lo = g.cpu().numpy(), lo.cpu().numpy()
        for i in range(2):
            Lr, _, gr = ref[i]
            assert relerr(g[i], gr) < 1e-4, (ks, rs, kg, rb, i)
            assert abs(lo[i] - Lr) < 1e-5 * Lr, (ks, rs, kg, rb, i)
            if base is not None:
                assert relerr(g[i], base[0][i]) < 1e-5, (ks, rs, kg, rb, i)
        if base is None:
            base = (g, lo)
    bank.close()


@pytest.mark.parametrize("Lx,Ly", [(300, 255), (257, 130)])
def test_band_rows_64_odd_sizes(pinned, Lx, Ly):
    import klt
    import torch
    print(f"case band rows 64 at {Lx}x{Ly}", flush=True)
    ctx = pinned
    n = 6
    rng = np.random.default_rng(Lx + Ly)
    T = (2 * n + 1) ** 2
    kr = (rng.standard_normal((8, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((8, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, 8))[::-1].astype(np.float32).copy()
    bank = ctx.kernels(kr, ki, sc, 2 * n + 1, 2 * n + 1)
    t = (rng.random((2, Ly, Lx)) > 0.6).astype(np.float32)
    m = np.clip(t + rng.uniform(-0.1, 0.1, t.shape), 0, 1).astype(np.float32)
    ctx.set_band_rows(64)
    img = ctx.socs_forward(bank, torch.from_numpy(m).cuda())
    ctx.sync()
    img = img.cpu().numpy()
    lo, g, _ = klt.socs_backward(ctx, bank, m, t, thr=0.02)
    for i in range(2):
        assert relerr(img[i], O.socs_image(m[i], kr, ki, sc, n, n)) < 1e-5
        Lr, _, gr = O.socs_loss_grad(m[i].astype(np.float64), t[i], kr, ki, sc, n, n, thr=0.02)
        assert relerr(g[i], gr) < 1e-4
        assert abs(lo[i] - Lr) < 1e-5 * Lr
    bank.close()


def test_oversized_factors_are_lowered(pinned):
    import torch
    ctx = pinned
    bank, _ = _bank(ctx, 41)
    rng = np.random.default_rng(10)
    dm = torch.from_numpy((rng.random((2, 256, 256)) > 0.5).astype(np.float32)).cuda()
    print("case lowered ks=64 at 256^2", flush=True)
    ctx.set_splits(2, 8, 3)
    a = ctx.socs_forward(bank, dm)
    ctx.sync()
    a = a.cpu().numpy()
    ctx.set_splits(64, 64, 64)
    b = ctx.socs_forward(bank, dm)
    ctx.sync()
    b = b.cpu().numpy()
    for i in range(2):
        assert relerr(b[i], a[i]) < 1e-6
    with pytest.raises(Exception):
        ctx.set_splits(-1, 0, 0)
    bank.close()
