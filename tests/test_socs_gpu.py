"""SOCS imaging on the GPU (klt_socs_forward) against the FP64 oracle.

Tolerance: image relative L2 <= 1e-5 against the oracle (FP32 device
arithmetic vs FP64), as the imaging target requires.  At full size the
size-independent properties of the SOCS sum are checked as well: clear field,
dark field, integer circular shift, linearity in the eigenvalues, batch ==
single image, host pointers == device pointers.
"""
import numpy as np
import pytest

import klt
import oracle as O
import workloads as W
from conftest import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _kernels(L, src_kind="Annular", nk=24, NA=1.35, defocus=0.0, nmed=1.0):
    src = O.source(src_kind, 51, inner=0.6, outer=0.9)
    sig = O.outer_sigma(src)
    Nx, Ny = O.sizing(NA, L, L, sig)
    t = O.tcc(src, sig, NA, defocus, L, L, Nx, Ny, nmed=nmed)
    w, v = O.eigh_topk(t, nk)
    kr, ki, sc = O.kernels_from_eig(v, w)
    return kr, ki, sc, Nx, Ny


@pytest.fixture(scope="module")
def bank29():
    return _kernels(W.PERIOD_29)


def _random_bank(rng, Nx, Ny, nk):
    T = (2 * Nx + 1) * (2 * Ny + 1)
    kr = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((nk, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.random(nk).astype(np.float32))[::-1].copy()
    return kr, ki, sc


SMALL = [(64, 64, 3, 3), (63, 65, 3, 2), (96, 80, 5, 6), (128, 128, 14, 14),
         (57, 57, 14, 14), (200, 120, 7, 4), (300, 260, 15, 15), (384, 256, 12, 9)]


@pytest.mark.parametrize("plain", [True, False], ids=["plain", "fused"])
@pytest.mark.parametrize("Lx,Ly,Nx,Ny", SMALL)
def test_small_random_kernels(ctx, Lx, Ly, Nx, Ny, plain):
    print(f"case Lx={Lx} Ly={Ly} Nx={Nx} Ny={Ny} plain={plain}", flush=True)
    rng = np.random.default_rng(Lx * 7 + Ly)
    kr, ki, sc = _random_bank(rng, Nx, Ny, 5)
    m = (rng.random((Ly, Lx)) > 0.5).astype(np.float32)
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    img = ctx.socs_forward(bank, m, plain=plain)
    assert relerr(img, O.socs_image(m, kr, ki, sc, Nx, Ny)) < TOL


# kernel half-sizes above 15 (the reference's own 33 x 33 sample has N = 16):
# odd and unequal periods, Ny != Nx.  "default" is the matrix-core path of
# klt_socs_forward (k_synth_wide), "plain" the FP32 single-image path.
# N 16-18: the single-block coarse stage; N 19-31: the four-launch one over
# device scratch (4N + 1 <= L at the limit: 125 <= 127)
LARGE_N = [(67, 70, 16, 16), (200, 75, 16, 18), (129, 100, 18, 17), (160, 160, 18, 18),
           (150, 140, 19, 20), (256, 200, 24, 22), (130, 127, 31, 31), (300, 260, 27, 31)]


@pytest.mark.parametrize("plain", [True, False], ids=["plain", "default"])
@pytest.mark.parametrize("Lx,Ly,Nx,Ny", LARGE_N)
def test_large_half_size_random_kernels(ctx, Lx, Ly, Nx, Ny, plain):
    print(f"case large N Lx={Lx} Ly={Ly} Nx={Nx} Ny={Ny} plain={plain}", flush=True)
    rng = np.random.default_rng(Lx * 11 + Ly + Nx)
    kr, ki, sc = _random_bank(rng, Nx, Ny, 4)
    m = (rng.random((Ly, Lx)) > 0.5).astype(np.float32)
    bank = ctx.kernels(kr, ki, sc, 2 * Nx + 1, 2 * Ny + 1)
    img = ctx.socs_forward(bank, m, plain=plain)
    assert relerr(img, O.socs_image(m, kr, ki, sc, Nx, Ny)) < TOL


def test_large_half_size_batch(ctx):
    """Three masks in one call at N = 25 (the coarse stage over device scratch
    takes the images in sub-batches): each image against the oracle."""
    print("case large N batch of 3, N=25", flush=True)
    rng = np.random.default_rng(25)
    kr, ki, sc = _random_bank(rng, 25, 25, 6)
    ms = np.stack([(rng.random((256, 256)) > 0.5).astype(np.float32) for _ in range(3)])
    bank = ctx.kernels(kr, ki, sc, 51, 51)
    img = ctx.socs_forward(bank, ms)
    for i in range(3):
        assert relerr(img[i], O.socs_image(ms[i], kr, ki, sc, 25, 25)) < TOL, i


def test_continuous_mask_fused(ctx):
    """Non-binary transmittance exercises the lo half of the split mask."""
    print("case continuous mask", flush=True)
    rng = np.random.default_rng(9)
    kr, ki, sc = _random_bank(rng, 14, 14, 8)
    m = W.relaxed(W.manhattan(512, 77), 5)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    assert relerr(ctx.socs_forward(bank, m), O.socs_image(m, kr, ki, sc, 14, 14)) < TOL


def test_config0_linespace(ctx):
    print("case config0 512 linespace", flush=True)
    kr, ki, sc, Nx, Ny = _kernels(512)
    assert (Nx, Ny) == (6, 6)
    m = np.zeros((512, 512), np.float32)
    O.lib().or_mask_linespace(m, 512, 512, 32, 32, 0)
    bank = ctx.kernels(kr, ki, sc, 13, 13)
    img = ctx.socs_forward(bank, m)
    assert relerr(img, O.socs_image(m, kr, ki, sc, Nx, Ny)) < TOL


def test_config1_manhattan_1024(ctx, bank29):
    print("case config1 1024 manhattan", flush=True)
    kr, ki, sc, Nx, Ny = bank29
    assert (Nx, Ny) == (14, 14)
    m = W.manhattan(1024, 1000)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    ref = O.socs_image(m, kr, ki, sc, Nx, Ny)
    for plain in (True, False):
        img = ctx.socs_forward(bank, m, plain=plain)
        assert relerr(img, ref) < TOL
        assert np.abs(img - ref).max() < 1e-4 * np.abs(ref).max()


def test_properties_full_size(ctx, bank29):
    import torch
    print("case properties 1024", flush=True)
    kr, ki, sc, Nx, Ny = bank29
    T = kr.shape[1]
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    ones = np.ones((1024, 1024), np.float32)
    clear = ctx.socs_forward(bank, ones)
    c = (T - 1) // 2
    expect = float(np.sum(sc.astype(np.float64) * (kr[:, c].astype(np.float64) ** 2
                                                    + ki[:, c].astype(np.float64) ** 2)))
    assert np.abs(clear - expect).max() < 1e-5 * expect
    dark = ctx.socs_forward(bank, np.zeros((1024, 1024), np.float32))
    assert np.abs(dark).max() == 0.0
    m = W.manhattan(1024, 1001)
    base = ctx.socs_forward(bank, m)
    sh = ctx.socs_forward(bank, np.ascontiguousarray(np.roll(m, (37, -91), (0, 1))))
    assert relerr(sh, np.roll(base, (37, -91), (0, 1))) < 1e-5
    bank2 = ctx.kernels(kr, ki, (2 * sc).astype(np.float32), 29, 29)
    assert relerr(ctx.socs_forward(bank2, m), 2 * base) < 1e-6
    # device pointers and batches
    masks = np.stack([W.manhattan(1024, 1002 + i) for i in range(4)])
    dm = torch.from_numpy(masks).cuda()
    dbank = ctx.kernels(torch.from_numpy(kr).cuda(), torch.from_numpy(ki).cuda(),
                        torch.from_numpy(sc).cuda(), 29, 29)
    dimg = ctx.socs_forward(dbank, dm)
    torch.cuda.synchronize()
    ctx.sync()
    out = dimg.cpu().numpy()
    for i in range(4):
        single = ctx.socs_forward(bank, masks[i])
        assert relerr(out[i], single) < 1e-6


def test_batch_uneven_small(ctx):
    print("case batch 256", flush=True)
    rng = np.random.default_rng(3)
    kr, ki, sc = _random_bank(rng, 14, 14, 24)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    masks = np.stack([W.manhattan(256, 2000 + i) for i in range(6)])
    out = ctx.socs_forward(bank, masks)
    for i in (0, 5):
        assert relerr(out[i], O.socs_image(masks[i], kr, ki, sc, 14, 14)) < TOL


def test_errors(ctx):
    rng = np.random.default_rng(4)
    kr, ki, sc = _random_bank(rng, 14, 14, 2)
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    with pytest.raises(klt.KltError) as e:
        ctx.socs_forward(bank, np.zeros((56, 56), np.float32))
    assert e.value.code == -2 and "FI" in e.value.msg
    with pytest.raises(klt.KltError) as e:
        ctx.kernels(np.zeros((1, 28 * 29), np.float32), np.zeros((1, 28 * 29), np.float32),
                    np.ones(1, np.float32), 28, 29)
    assert e.value.code == -1 and "odd" in e.value.msg
