"""The resist step on the GPU (klt_resist_levels: the reference's 7-level view
of the image against TARGET_INTENSITY, klitho_tcc.cpp:442,547,
common/file_io.cpp:25-80) against the oracle's classes, bit for bit: random
values with the class boundaries, NaN and infinities, odd sizes (the scalar
tail of the kernel), min == max, and a 1024^2 SOCS image batch on device
pointers (level 6 = the printed resist, I >= 0.225)."""
import numpy as np
import pytest

import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,lo,hi", [((3, 37, 23), 0.0, 0.225), ((2, 64, 64), -1.5, 2.0),
                                         ((1, 5, 9), 3.0, 3.0)])
def test_resist_levels_vs_oracle(ctx, shape, lo, hi):
    print(f"case resist levels {shape} [{lo}, {hi}]", flush=True)
    rng = np.random.default_rng(sum(shape))
    d = rng.uniform(lo - 0.3 * (hi - lo + 1), hi + 0.3 * (hi - lo + 1), shape).astype(np.float32)
    f = d.reshape(-1)
    edges = np.float32(lo) + (np.arange(7, dtype=np.float32) / np.float32(6)) * np.float32(hi - lo)
    f[:7] = edges
    f[7:14] = np.nextafter(edges, np.float32(-np.inf))
    f[14:17] = [np.nan, np.inf, -np.inf]
    got = ctx.resist_levels(d, lo, hi)
    assert np.array_equal(got, O.resist_levels(d, lo, hi))


def test_resist_levels_socs_batch_device(ctx):
    import torch
    print("case resist levels on 4 x 1024^2 SOCS images (device)", flush=True)
    rng = np.random.default_rng(3)
    T = 29 * 29
    kr = (rng.standard_normal((8, T)) / np.sqrt(T)).astype(np.float32)
    ki = (rng.standard_normal((8, T)) / np.sqrt(T)).astype(np.float32)
    sc = np.sort(rng.uniform(0.1, 2.0, 8))[::-1].astype(np.float32).copy()
    bank = ctx.kernels(kr, ki, sc, 29, 29)
    m = torch.from_numpy(W.manhattan_batch(1024, 4, 40)).cuda()
    img = ctx.socs_forward(bank, m)
    h = img.cpu().numpy()
    # the reference's threshold, and one inside this image's range (random
    # kernels give a dim image) so that every class occurs
    for vmax in (0.225, float(np.percentile(h, 70))):
        lev = ctx.resist_levels(img, 0.0, vmax)
        torch.cuda.synchronize()
        ref = O.resist_levels(h, 0.0, vmax)
        assert np.array_equal(lev.cpu().numpy(), ref), vmax
    assert (ref == 6).any() and (ref < 6).any()
    bank.close()
