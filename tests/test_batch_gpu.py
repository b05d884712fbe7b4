"""Mixed-size, multi-set batches (klt_socs_forward_batch).

Each image of a batch must equal the same mask imaged alone through
klt_socs_forward with the same kernel set: the same per-image arithmetic runs
in both, up to the order of the FP32 band sums when the single call splits
its band pass over column ranges. The batch tests therefore check:
  * agreement with the single-size calls (relative L2 <= 1e-6), and against
    the FP64 oracle (image relative L2 <= 1e-5);
  * graph replays with the program's own eager work between them: the same
    buffers refilled with new masks give the new images on every replay;
  * the per-image completion times of the last batch;
  * errors: host pointers, kernel sets of different sizes, 4N+1 > L.
"""
import numpy as np
import pytest

import klt
import oracle as O
from conftest import relerr

pytestmark = pytest.mark.gpu
N = 6


def _bank(ctx, seed, nk=4, n=N):
    rng = np.random.default_rng(seed)
    T = (2 * n + 1) ** 2
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.1, 4.0, nk))[::-1].astype(np.float32).copy()
    return ctx.kernels(kr, ki, sc, 2 * n + 1, 2 * n + 1), (kr, ki, sc)


def _masks(torch, sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for (ly, lx) in sizes:
        m = (rng.random((ly, lx)) > 0.55).astype(np.float32)
        out.append(torch.from_numpy(m).cuda())
    return out


SIZES = [(64, 64), (96, 128), (128, 96), (160, 160), (75, 100), (256, 256)]


def test_batch_equals_single_calls_and_oracle(ctx):
    import torch
    print("case batch sizes", SIZES, flush=True)
    banks = [_bank(ctx, 10 + s) for s in range(3)]
    masks = _masks(torch, SIZES, 1)
    imgs = ctx.socs_forward_batch([b for b, _ in banks], masks)
    torch.cuda.synchronize()
    ctx.sync()
    for m, mk in enumerate(masks):
        for s, (bank, (kr, ki, sc)) in enumerate(banks):
            got = imgs[m * 3 + s].cpu().numpy()
            single = ctx.socs_forward(bank, mk.clone())
            ctx.sync()
            assert relerr(got, single.cpu().numpy()) < 1e-6, (SIZES[m], s)
            if m in (0, 4):  # one even, one odd size against FP64
                ref = O.socs_image(mk.cpu().numpy(), kr, ki, sc, N, N)
                assert relerr(got, ref) < 1e-5, (SIZES[m], s)


def test_batch_graph_replays_follow_new_data(ctx):
    import torch
    print("case batch replays", flush=True)
    banks = [_bank(ctx, 20 + s) for s in range(2)]
    sizes = [(128, 128), (96, 64), (64, 96)]
    masks = _masks(torch, sizes, 2)
    imgs = [torch.empty_like(m) for m in masks for _ in banks]
    for rep in range(4):
        # the program's own eager work between replays: refill the masks
        rng = np.random.default_rng(100 + rep)
        for mk in masks:
            mk.copy_(torch.from_numpy(
                (rng.random(tuple(mk.shape)) > 0.5).astype(np.float32)).cuda())
        torch.cuda.synchronize()
        ctx.socs_forward_batch([b for b, _ in banks], masks, imgs)
        ctx.sync()
        for m, mk in enumerate(masks):
            for s, (bank, _) in enumerate(banks):
                single = ctx.socs_forward(bank, mk.clone())
                ctx.sync()
                assert relerr(imgs[m * 2 + s].cpu().numpy(), single.cpu().numpy()) < 1e-6, rep
        lat = ctx.batch_latency()
        assert lat.shape == (len(masks) * len(banks),)
        assert np.all(np.isfinite(lat)) and np.all(lat > 0) and lat.max() < 1e6


def test_batch_errors(ctx):
    import ctypes as C
    import torch
    print("case batch errors", flush=True)
    b6, _ = _bank(ctx, 30)
    b5, _ = _bank(ctx, 31, n=5)
    mk = _masks(torch, [(64, 64)], 3)
    with pytest.raises(klt.KltError) as e:
        ctx.socs_forward_batch([b6, b5], mk)
    assert e.value.code == -1
    small = _masks(torch, [(20, 64)], 4)
    with pytest.raises(klt.KltError) as e:
        ctx.socs_forward_batch([b6], small)
    assert e.value.code == -2 and "FI" in e.value.msg
    mp = (C.c_void_p * 1)(mk[0].data_ptr())
    lx = (C.c_int * 1)(64)
    out = torch.empty_like(mk[0])
    ip = (C.c_void_p * 1)(out.data_ptr())
    bp = (C.c_void_p * 1)(b6.handle.value)
    rc = klt.lib().klt_socs_forward_batch(ctx.handle, 1, mp, lx, lx, 1, bp, ip, klt.KLT_HOST_PTRS)
    assert rc == -1


def test_batch_captured_by_caller_survives_later_batches(ctx):
    """A batch launched into the caller's own graph capture (torch.cuda.graph)
    keeps its buffers while later distinct batches fill and evict the
    library's graph cache; each replay of the caller's graph, with the
    program's eager refills between replays, equals the single calls."""
    import torch
    print("case batch captured by caller", flush=True)
    banks = [_bank(ctx, 40 + s) for s in range(2)]
    bl = [b for b, _ in banks]
    sizes = [(128, 128), (96, 64)]
    masks = _masks(torch, sizes, 5)
    imgs = [torch.empty_like(m) for m in masks for _ in banks]
    s = torch.cuda.Stream()
    ctx.set_stream(s.cuda_stream)
    try:
        with torch.cuda.stream(s):
            ctx.socs_forward_batch(bl, masks, imgs)  # first call: outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            ctx.socs_forward_batch(bl, masks, imgs)
        torch.cuda.synchronize()
        # five other batches: more distinct entries than the cache keeps
        for j in range(5):
            other = _masks(torch, [(64 + 16 * j, 80)], 50 + j)
            with torch.cuda.stream(s):
                ctx.socs_forward_batch(bl, other)
        torch.cuda.synchronize()
        for rep in range(3):
            rng = np.random.default_rng(200 + rep)
            for mk in masks:
                mk.copy_(torch.from_numpy(
                    (rng.random(tuple(mk.shape)) > 0.5).astype(np.float32)).cuda())
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for m, mk in enumerate(masks):
                for si, bank in enumerate(bl):
                    with torch.cuda.stream(s):
                        single = ctx.socs_forward(bank, mk.clone())
                    torch.cuda.synchronize()
                    assert relerr(imgs[m * 2 + si].cpu().numpy(),
                                  single.cpu().numpy()) < 1e-6, (rep, m, si)
    finally:
        torch.cuda.synchronize()
        ctx.set_stream(None)
        ctx.sync()


WIDE_SIZES = [(256, 256), (384, 200), (130, 300), (512, 512)]


def test_batch_wide_half_size(ctx):
    """Kernel half-size 16 (the wide passes: 64 twiddle rows, 128-point coarse
    stage, K = 128 synthesis) in a mixed batch of two kernel sets: every image
    against the oracle (1e-5) and the single-size calls (1e-5: the single calls
    may split the band pass over column ranges and the coarse stage over kernel
    groups), then one graph replay over refilled buffers."""
    import torch
    print("case batch wide N=16 sizes", WIDE_SIZES, flush=True)
    banks = [_bank(ctx, 40 + s, nk=5, n=16) for s in range(2)]
    masks = _masks(torch, WIDE_SIZES, 7)
    imgs = None
    for rep in range(2):
        if rep:
            rng = np.random.default_rng(8)
            for mk in masks:
                mk.copy_(torch.from_numpy((rng.random(tuple(mk.shape)) > 0.4).astype(np.float32)))
        # (the second call, same buffers, replays the recorded graph)
        imgs = ctx.socs_forward_batch([b for b, _ in banks], masks, imgs)
        torch.cuda.synchronize()
        ctx.sync()
        for m, mk in enumerate(masks):
            mh = mk.cpu().numpy()
            for s, (bank, (kr, ki, sc)) in enumerate(banks):
                got = imgs[m * 2 + s].cpu().numpy()
                single = ctx.socs_forward(bank, mk.clone())
                ctx.sync()
                assert relerr(got, single.cpu().numpy()) < 1e-5, (rep, WIDE_SIZES[m], s)
                if s == 0:
                    assert relerr(got, O.socs_image(mh, kr, ki, sc, 16, 16)) < 1e-5, (rep, m)
    lat = ctx.batch_latency()
    assert lat.shape == (len(WIDE_SIZES) * 2,) and np.all(lat > 0)
    for b, _ in banks:
        b.close()
