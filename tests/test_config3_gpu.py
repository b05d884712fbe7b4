"""Uneven batch at its own shape (klt_socs_forward_batch) against the FP64 oracle.

Workload: seeded Manhattan masks with sides from {256, 384, 512, 768, 1024}
(two of each, in shuffled order) under three kernel sets of 24 kernels of
29 x 29 (annular 0.6 / 0.9 source, NA 1.35 in water, period 1088 nm, defocus
-25 / 0 / +25 nm; TCC and kernels from the library), all images of the batch
in one call (K-Litho-SOCS/klitho_socs.cpp:196-241 per image).

* one image per side and per set against the oracle's SOCS image of the same
  mask and kernels: relative L2 <= 1e-5;
* a replay of the captured graph over the same buffers refilled with new masks:
  every image equals the single-size call on the new mask (relative L2 <= 1e-6),
  one image per side (set 0) against the oracle.
"""
import numpy as np
import pytest

import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu

SIDES = [256, 384, 512, 768, 1024]
DEFOCUS_NM = [-25, 0, 25]


@pytest.fixture(scope="module")
def sets(ctx):
    src = O.source("Annular", 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    out = []
    for df_nm in DEFOCUS_NM:
        t = ctx.tcc_build(src, sig, W.NA, df_nm * W.NA ** 2 / W.LAMBDA, W.PERIOD_29,
                          W.PERIOD_29, 14, 14, precision=0, nmed=1.44)
        kr, ki, sc = ctx.calc_kernels(t, 24)
        out.append((ctx.kernels(kr, ki, sc, 29, 29), (kr, ki, sc)))
    yield out
    for b, _ in out:
        b.close()


def test_config3_batch_vs_oracle_and_replay(ctx, sets):
    import torch
    rng = np.random.default_rng(11)
    sides = [s for s in SIDES for _ in range(2)]
    rng.shuffle(sides)
    print(f"case config3 batch sides {sides} x {len(sets)} sets", flush=True)
    masks = [torch.from_numpy(W.manhattan(n, 6000 + i)).cuda() for i, n in enumerate(sides)]
    banks = [b for b, _ in sets]
    nb = len(banks)
    images = [torch.empty_like(m) for m in masks for _ in banks]
    ctx.socs_forward_batch(banks, masks, images)  # captures the graph
    torch.cuda.synchronize()
    ctx.sync()
    first = {}
    for i, n in enumerate(sides):
        first.setdefault(n, i)
    for n, i in first.items():
        m = masks[i].cpu().numpy()
        for s, (_, (kr, ki, sc)) in enumerate(sets):
            got = images[i * nb + s].cpu().numpy()
            assert relerr(got, O.socs_image(m, kr, ki, sc, 14, 14)) < 1e-5, (n, s)
    # refill the same buffers with new masks; the call replays the graph
    for i, n in enumerate(sides):
        masks[i].copy_(torch.from_numpy(W.manhattan(n, 7000 + i)).cuda())
    torch.cuda.synchronize()
    ctx.socs_forward_batch(banks, masks, images)
    torch.cuda.synchronize()
    ctx.sync()
    for i, n in enumerate(sides):
        for s, bank in enumerate(banks):
            single = ctx.socs_forward(bank, masks[i].clone())
            ctx.sync()
            torch.cuda.synchronize()
            assert relerr(images[i * nb + s].cpu().numpy(), single.cpu().numpy()) < 1e-6, (n, s)
    kr, ki, sc = sets[0][1]
    for n, i in first.items():
        ref = O.socs_image(masks[i].cpu().numpy(), kr, ki, sc, 14, 14)
        assert relerr(images[i * nb].cpu().numpy(), ref) < 1e-5, n
    lat = ctx.batch_latency()
    assert lat.shape == (len(images),) and np.all(lat > 0)
