"""klt_socs_forward_noalloc (no device allocation; every table and intermediate
in the library's fixed per-GPU scratch array) against klt_socs_forward, the
entry point whose kernels it shares, called through the C ABI:

* two calls in a row, the second with a smaller input (fewer masks, another
  geometry, so every table is regenerated in the shared array): each image
  within relative L2 1e-6 of klt_socs_forward's (the same FP32 arithmetic; the
  tables are generated on the device, and the band sums may be split
  differently), and one image of each call within 1e-5 of the FP64 oracle;
* a batch larger than one sub-batch of the array (1024^2 masks);
* errors: sides above KLT_NOALLOC_MAX_SIDE, kernel half-size above 15.
"""
import ctypes as C

import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu


def _bank(ctx, seed, n, nk):
    rng = np.random.default_rng(seed)
    T = (2 * n + 1) ** 2
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.1, 4.0, nk))[::-1].astype(np.float32).copy()
    return ctx.kernels(kr, ki, sc, 2 * n + 1, 2 * n + 1), (kr, ki, sc)


def _noalloc(ctx, bank, masks, images):
    import torch
    B, Ly, Lx = masks.shape
    torch.cuda.synchronize()
    rc = klt.lib().klt_socs_forward_noalloc(ctx.handle, bank.handle, C.c_void_p(masks.data_ptr()),
                                            B, Lx, Ly, C.c_void_p(images.data_ptr()))
    klt.check(rc)
    ctx.sync()


def _case(ctx, torch, L_x, L_y, B, n, nk, seed):
    bank, (kr, ki, sc) = _bank(ctx, seed, n, nk)
    hm = np.stack([W.manhattan(max(L_x, L_y), seed + i)[:L_y, :L_x] for i in range(B)])
    masks = torch.from_numpy(np.ascontiguousarray(hm)).cuda()
    got = torch.full_like(masks, -1.0)
    _noalloc(ctx, bank, masks, got)
    ref = ctx.socs_forward(bank, masks.clone())
    torch.cuda.synchronize()
    ctx.sync()
    g, r = got.cpu().numpy(), ref.cpu().numpy()
    for i in range(B):
        assert relerr(g[i], r[i]) < 1e-6, (L_x, L_y, i)
    assert relerr(g[0], O.socs_image(hm[0], kr, ki, sc, n, n)) < 1e-5
    bank.close()


def test_noalloc_matches_forward_two_calls(ctx):
    import torch
    print("case noalloc 1024^2 x 6 (N=14, K=24), then 300 x 260 x 2 (N=9, K=5)", flush=True)
    _case(ctx, torch, 1024, 1024, 6, 14, 24, 500)
    _case(ctx, torch, 300, 260, 2, 9, 5, 600)


def test_noalloc_sub_batches(ctx):
    """1024^2 images take ~44 KB of intermediates each: 700 masks run as two
    sub-batches of the 28 MiB work area."""
    import torch
    print("case noalloc 700 x 1024^2 (two sub-batches)", flush=True)
    bank, _ = _bank(ctx, 700, 14, 8)
    hm = np.stack([W.manhattan(1024, 900 + i) for i in range(4)])
    masks = torch.from_numpy(hm).cuda().repeat(175, 1, 1).contiguous()
    got = torch.empty_like(masks)
    _noalloc(ctx, bank, masks, got)
    ref = ctx.socs_forward(bank, masks)
    torch.cuda.synchronize()
    ctx.sync()
    for i in (0, 1, 350, 661, 662, 699):
        assert relerr(got[i].cpu().numpy(), ref[i].cpu().numpy()) < 1e-6, i
    bank.close()


def test_noalloc_errors(ctx):
    import torch
    print("case noalloc errors", flush=True)
    bank, _ = _bank(ctx, 800, 6, 2)
    m = torch.zeros((1, 64, 4100), dtype=torch.float32, device="cuda")
    rc = klt.lib().klt_socs_forward_noalloc(ctx.handle, bank.handle, C.c_void_p(m.data_ptr()), 1,
                                            4100, 64, C.c_void_p(m.data_ptr()))
    assert rc == -1  # KLT_ERR_ARG
    assert b"4096" in klt.lib().klt_last_error()
    big, _ = _bank(ctx, 801, 16, 1)
    m2 = torch.zeros((1, 128, 128), dtype=torch.float32, device="cuda")
    rc = klt.lib().klt_socs_forward_noalloc(ctx.handle, big.handle, C.c_void_p(m2.data_ptr()), 1,
                                            128, 128, C.c_void_p(m2.data_ptr()))
    assert rc == -1
    bank.close()
    big.close()
