"""Batches larger than one launch: klt_socs_forward runs batches of more than
1024 masks and klt_socs_backward batches of more than 512 as consecutive
sub-batches that share the context's scratch.

Checks, at 64 x 64 with a random kernel bank (N = 3, 4 kernels):
  * 1100 masks forward: each image equals the same mask's image from calls of
    100 masks (relative L2 1e-6 per image: the two shapes may be tuned to
    different split factors, which reorders FP32 sums), and the images on both
    sides of the sub-batch boundary (0, 1023, 1024, 1099) match the FP64
    oracle to 1e-5;
  * 520 masks backward: gradients 1e-4 and losses 1e-5 against the FP64
    oracle at masks 0, 511, 512, 519 for both calls, and 1e-5 per image
    between them (the fused synthesis scales each block's resist-band
    accumulators by a power of two from the block's largest entry before
    their fp16 split, so other row splits round them differently; losses
    1e-6).
"""
import numpy as np
import pytest

import klt
import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu
N, NK, L = 3, 4, 64


def _bank(ctx, seed=17):
    rng = np.random.default_rng(seed)
    T = (2 * N + 1) ** 2
    kr = (rng.standard_normal((NK, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((NK, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.2, 3.0, NK))[::-1].astype(np.float32).copy()
    return ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1), (kr, ki, sc)


def _per_image(a, b):
    return max(relerr(a[i], b[i]) for i in range(a.shape[0]))


def test_forward_over_1024_masks(ctx):
    print("case forward 1100 x 64^2", flush=True)
    bank, (kr, ki, sc) = _bank(ctx)
    rng = np.random.default_rng(3)
    masks = (rng.random((1100, L, L)) > 0.5).astype(np.float32)
    full = ctx.socs_forward(bank, masks)
    parts = np.concatenate([ctx.socs_forward(bank, masks[b:b + 100])
                            for b in range(0, 1100, 100)])
    assert _per_image(full, parts) < 1e-6
    for i in (0, 1023, 1024, 1099):
        assert relerr(full[i], O.socs_image(masks[i], kr, ki, sc, N, N)) < 1e-5, i
    bank.close()


def test_backward_over_512_masks(ctx):
    print("case backward 520 x 64^2", flush=True)
    bank, (kr, ki, sc) = _bank(ctx, seed=23)
    rng = np.random.default_rng(4)
    tgts = (rng.random((520, L, L)) > 0.6).astype(np.float32)
    masks = np.clip(tgts + rng.uniform(-0.1, 0.1, tgts.shape), 0, 1).astype(np.float32)
    l_full, g_full, _ = klt.socs_backward(ctx, bank, masks, tgts, thr=0.02)
    outs = [klt.socs_backward(ctx, bank, masks[b:b + 100], tgts[b:b + 100], thr=0.02)
            for b in range(0, 520, 100)]
    l_parts = np.concatenate([o[0] for o in outs])
    g_parts = np.concatenate([o[1] for o in outs])
    for i in (0, 511, 512, 519):
        L0, _, g = O.socs_loss_grad(masks[i].astype(np.float64), tgts[i], kr, ki, sc, N, N,
                                    thr=0.02)
        print(f"mask {i}: gradient rel-L2 vs FP64 {relerr(g_full[i], g):.2e} (520-mask call), "
              f"{relerr(g_parts[i], g):.2e} (100-mask calls)", flush=True)
        assert relerr(g_full[i], g) < 1e-4, i
        assert relerr(g_parts[i], g) < 1e-4, i
        assert abs(l_full[i] - L0) < 1e-5 * L0, i
    d = _per_image(g_full, g_parts)
    print(f"largest per-image gradient difference between the calls {d:.2e}", flush=True)
    assert d < 1e-5
    assert np.max(np.abs(l_full - l_parts) / np.abs(l_parts)) < 1e-6
    bank.close()
