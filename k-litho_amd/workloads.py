"""Synthetic inputs of the benchmark workloads (seeded, reproducible).

* manhattan(L, seed): L x L float32 binary mask of non-overlapping
  axis-aligned rectangles on a 4 px grid.  At L = 1024: 60..150 rectangles,
  sides uniform in 16..128 px.  Other sizes scale the side range by L/1024
  and the count range by (L/1024)^2 (minimum side 4 px, minimum count 4).
* relaxed(mask, seed): clip(mask + U(-0.1, 0.1), 0, 1) (adjoint workload).
* blurred(src): a source grid convolved with a Gaussian of 1.5 cells along
  both axes (truncated at 4 sigma, zero outside the grid), normalised to sum 1:
  applied to the 101 x 101 binary dipole (sigma_c 0.7, radius 0.2, on x) it is
  the source of the T = 1225 TCC precision study.
* Optics of the workloads: lambda 193 nm, NA 1.35, 51 x 51 annular source
  sigma 0.6 / 0.9, and the TCC periods that give the stated kernel supports
  (1088 nm -> 29 x 29 kernels, 1280 nm -> 35 x 35).
"""
from __future__ import annotations

import numpy as np

LAMBDA = 193.0
NA = 1.35
SRC_SIZE = 51
SIGMA_IN, SIGMA_OUT = 0.6, 0.9
# period whose sizing formula floor(NA*L*(1+sigma_out)/lambda) gives Nx = 14
PERIOD_29 = 1088
# ... and Nx = 17 (35 x 35, T = 1225) for the dipole study
PERIOD_35 = 1280


def manhattan(L: int, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    scale = L / 1024.0
    lo_n = max(4, int(round(60 * scale * scale)))
    hi_n = max(lo_n, int(round(150 * scale * scale)))
    lo_s = max(4, int(round(16 * scale / 4)) * 4)
    hi_s = max(lo_s, int(round(128 * scale / 4)) * 4)
    target = int(rng.integers(lo_n, hi_n + 1))
    occ = np.zeros((L // 4 + 1, L // 4 + 1), bool)  # 4 px cells
    m = np.zeros((L, L), np.float32)
    placed, tries = 0, 0
    while placed < target and tries < 50 * target:
        tries += 1
        w = int(rng.integers(lo_s // 4, hi_s // 4 + 1))
        h = int(rng.integers(lo_s // 4, hi_s // 4 + 1))
        cx = int(rng.integers(0, L // 4 - w + 1))
        cy = int(rng.integers(0, L // 4 - h + 1))
        if occ[cy:cy + h, cx:cx + w].any():
            continue
        occ[cy:cy + h, cx:cx + w] = True
        m[4 * cy:4 * (cy + h), 4 * cx:4 * (cx + w)] = 1.0
        placed += 1
    return m


def relaxed(mask: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = rng.uniform(-0.1, 0.1, size=mask.shape)
    return np.clip(mask + noise, 0.0, 1.0).astype(np.float32)


def blurred(src: np.ndarray, sigma_cells: float = 1.5) -> np.ndarray:
    r = np.arange(-int(4 * sigma_cells), int(4 * sigma_cells) + 1)
    k = np.exp(-0.5 * (r / sigma_cells) ** 2)
    k /= k.sum()
    b = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 0,
                            np.asarray(src, np.float64))
    b = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), 1, b)
    return np.ascontiguousarray((b / b.sum()).astype(np.float32))


def manhattan_batch(L: int, n: int, seed0: int = 1000) -> np.ndarray:
    return np.stack([manhattan(L, seed0 + i) for i in range(n)])
