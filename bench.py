"""Benchmark of the SOCS imaging hot path on one MI355X.

Headline workload (1024^2 SOCS imaging, K = 24): seeded 1024 x 1024 Manhattan
masks and the 24 kernels of 29 x 29 that the library itself derives on the
GPU (TCC of the 51 x 51 annular source sigma 0.6/0.9 at NA 1.35, period
1088 nm, then the FP64 eigensolver).  A step is one klt_socs_forward call
over a batch of masks already resident in HBM; `value` is images per second
over the timed steps.

Also reported (same JSON line):
  plain        the FP32 single-image path (klt_socs_image, one call per image)
  adjoint      klt_socs_backward on 16 (and 128) relaxed 1024^2 masks (mask
               gradients/s, with its own roofline object), klt_socs_vjp (the
               product with a given dL/dI) on the same masks
               and klt_abbe_source_grad (source gradient over the 51^2 grid)
  tcc_eig      TCC T = 1225 (blurred dipole, period 1280, 5 defocus values),
               split-FP16 MFMA Gram vs FP64 Gram, top-32 FP64 eigensolve (one
               matrix per call, and the five in one batched call)
  batch_uneven 64 masks of sides {256..1024} x 3 kernel sets (defocus -25/0/25
               nm) = 192 images per klt_socs_forward_batch graph launch, with
               p50/p99 per-image latency and the same work as grouped calls
  wide         kernel half-sizes above 15: the reference's own test_tcc.sh
               imaging (2048^2, 3 kernels of 33 x 33) and N = 31 (24 kernels of
               63 x 63 at 1024^2), forward images/s and adjoint mask grads/s
  kernels      per-kernel HIP-event times of one step; `roofline` is the
               dominant kernel's algorithmic HBM bytes over its duration
  cpu_baseline the reference's FFTW3 path (FT_r2c + calcSOCS + FI) restated
               in C++ on the FFTW3 interface of the host's MKL, 1 core and all
               cores, calcTCC and zheevr_ top-32 at T = 1225, on bounded
               samples, with the GPU image's (and a 512^2 adjoint sample's)
               relative L2 error vs the FP64 oracle

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--batch 512]
    python bench.py --copy-calibration   # also times a 1 GiB torch copy
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "k-litho_amd"))

import klt  # noqa: E402
import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_F16_PEAK_TFLOPS = 2500.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
L = 1024
NK = 24
KSIZE = 29
NMED = 1.44  # water immersion index: NA 1.35 needs n >= NA; at zero defocus it has no effect


def log(*a):
    print(*a, file=sys.stderr, flush=True)


COUNTERS = os.path.join(ROOT, "profiles", "socs_forward_1024_hbm_counters.json")


def measured_traffic(kernel, images_per_launch):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (copy-calibrated), scaled from their batch to this one;
    None when no measurement for this kernel and size is on file."""
    try:
        with open(COUNTERS) as f:
            c = json.load(f)
        k = c["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    if c.get("L") != L:
        return None
    return k["bytes_per_launch"] * images_per_launch / c["batch"]


class Timer:
    """HIP-event timing on one stream."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def __call__(self, fn, reps=1):
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        t.cuda.synchronize()
        e0.record(self.stream)
        for _ in range(reps):
            fn()
        e1.record(self.stream)
        t.cuda.synchronize()
        return e0.elapsed_time(e1) / reps  # ms


def rel_l2(got, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(np.asarray(got, np.float64) - ref) / np.linalg.norm(ref))


def annular_source():
    return klt.source_create(W.SRC_SIZE, "Annular", W.SIGMA_IN, W.SIGMA_OUT)


def gpu_kernels(ctx):
    """The 24 kernels of 29 x 29 of the 1024^2 workload, built on the GPU."""
    src = annular_source()
    sig = min(klt.outer_sigma(src), 1.0)
    Nx, Ny = klt.tcc_sizing(W.NA, W.PERIOD_29, W.PERIOD_29, sig)
    assert (Nx, Ny) == (14, 14)
    tcc = ctx.tcc_build(src, sig, W.NA, 0.0, W.PERIOD_29, W.PERIOD_29, Nx, Ny, precision=0,
                        nmed=NMED)
    kr, ki, sc = ctx.calc_kernels(tcc, NK)
    return src, kr, ki, sc


def bench_batch_uneven(ctx, timer, torch, src):
    """64 seeded Manhattan masks with sides from {256, 384, 512, 768, 1024}, each
    under 3 kernel sets (annular source, defocus -25 / 0 / +25 nm, K = 24,
    29 x 29): 192 images per klt_socs_forward_batch call (one graph launch)."""
    rng = np.random.default_rng(3)
    sides = [int(v) for v in rng.choice([256, 384, 512, 768, 1024], 64)]
    masks = [torch.from_numpy(W.manhattan(n, 2000 + i)).cuda() for i, n in enumerate(sides)]
    sig = min(klt.outer_sigma(src), 1.0)
    banks, defocus = [], [-25, 0, 25]
    for df_nm in defocus:
        tcc = ctx.tcc_build(src, sig, W.NA, df_nm * W.NA ** 2 / W.LAMBDA, W.PERIOD_29,
                            W.PERIOD_29, 14, 14, precision=0, nmed=NMED)
        kr, ki, sc = ctx.calc_kernels(tcc, NK)
        banks.append(ctx.kernels(kr, ki, sc, KSIZE, KSIZE))
    images = [torch.empty_like(m) for m in masks for _ in banks]
    for _ in range(3):  # the first call captures the graph
        ctx.socs_forward_batch(banks, masks, images)
    reps = 20
    ms = timer(lambda: ctx.socs_forward_batch(banks, masks, images), reps)
    lat = ctx.batch_latency()
    # the same images as per-size, per-set klt_socs_forward calls
    groups = {}
    for i, n in enumerate(sides):
        groups.setdefault(n, []).append(i)
    stacked = {n: torch.stack([masks[i] for i in idx]) for n, idx in groups.items()}
    outs = {n: torch.empty_like(t) for n, t in stacked.items()}

    def grouped():
        for b in banks:
            for n, t in stacked.items():
                ctx.socs_forward(b, t, outs[n])

    grouped()
    gms = timer(grouped, 5)
    err = 0.0
    for i in groups[sides[0]][:2]:
        for s, b in enumerate(banks):
            one = ctx.socs_forward(b, masks[i].clone())
            ctx.sync()
            err = max(err, rel_l2(images[i * 3 + s].cpu().numpy(), one.cpu().numpy()))
    nbytes = sum(4 * n * n for n in sides) * (1 + len(banks))
    for b in banks:
        b.close()
    return {"masks": len(sides), "kernel_sets": len(banks), "images": len(images),
            "defocus_nm": defocus,
            "sides": {str(n): len(groups.get(n, [])) for n in (256, 384, 512, 768, 1024)},
            "ms_per_batch": ms, "images_per_s": len(images) / (ms / 1e3),
            "p50_image_latency_us": float(np.percentile(lat, 50)),
            "p99_image_latency_us": float(np.percentile(lat, 99)),
            "algorithmic_hbm_bytes": nbytes, "hbm_gbs": nbytes / (ms / 1e3) / 1e9,
            "ms_per_batch_grouped_calls": gms,
            "max_rel_l2_vs_single_calls": err,
            "note": "one mask read per mask (its band is shared by the 3 sets), one image "
                    "write per image; latency = device wall clock from the batch start to "
                    "the last strip of each image"}


def blurred_dipole():
    """101 x 101 binary dipole (sigma_c 0.7, radius 0.2, on x) blurred by a
    Gaussian of 1.5 cells (truncated at 4 sigma), normalised."""
    return W.blurred(klt.source_create(101, "Dipole", radius=0.2, offset=0.7, on_x_axis=1))


def bench_tcc_eig(ctx, timer, torch):
    """TCC of the blurred dipole at T = 1225 over 5 defocus values, on the device:
    split-FP16 MFMA Gram vs FP64 Gram, per-kernel times, then the top-32 FP64
    eigensolve."""
    src = blurred_dipole()
    sig = min(klt.outer_sigma(src), 1.0)
    Nx, Ny = klt.tcc_sizing(W.NA, W.PERIOD_35, W.PERIOD_35, sig)
    T = (2 * Nx + 1) * (2 * Ny + 1)
    sh = 50
    yy, xx = np.mgrid[-sh:sh + 1, -sh:sh + 1]
    used = int(np.count_nonzero((src != 0) & (xx * xx + yy * yy <= sh * sh)))
    out = {"T": T, "source_points": used, "defocus_nm": [-50, -25, 0, 25, 50]}
    dsrc = torch.from_numpy(src).cuda()
    t64 = torch.empty((T, T, 2), dtype=torch.float64, device="cuda")
    t16 = torch.empty_like(t64)
    t5 = torch.empty((len(out["defocus_nm"]), T, T, 2), dtype=torch.float64, device="cuda")
    rows, gram_ms, eig_kernels = [], [], {}
    for di, df_nm in enumerate(out["defocus_nm"]):
        df = df_nm * W.NA ** 2 / W.LAMBDA

        def build(prec, dst):
            return ctx.tcc_build(dsrc, sig, W.NA, df, W.PERIOD_35, W.PERIOD_35, Nx, Ny, prec,
                                 NMED, out=dst)

        build(0, t64)
        build(1, t16)
        ctx.sync()
        t5[di].copy_(t64)
        a = t64.cpu().numpy()
        b = t16.cpu().numpy()
        a = a[..., 0] + 1j * a[..., 1]
        b = b[..., 0] + 1j * b[..., 1]
        ms16 = timer(lambda: build(1, t16), 3)
        ms64 = timer(lambda: build(0, t64), 3)
        ctx.set_profiling(True)
        build(1, t16)
        rep = ctx.profile_report()
        ctx.set_profiling(False)
        gram_ms.append(rep["k_gram_split16"][1] + rep["k_gram_reduce"][1])
        w64, _ = ctx.eig_topk(t64, 32, psd=True)
        w16, _ = ctx.eig_topk(t16, 32, psd=True)
        mse = timer(lambda: ctx.eig_topk(t64, 32, psd=True), 3)
        mshh = timer(lambda: ctx.eig_topk(t64, 32), 1)
        if df_nm == 0:
            for name, flag in (("subspace", True), ("householder", False)):
                ctx.set_profiling(True)
                ctx.eig_topk(t64, 32, psd=flag)
                eig_kernels[name] = {k: {"calls": v[0], "ms": v[1]}
                                     for k, v in ctx.profile_report().items()}
                ctx.set_profiling(False)
        w64 = np.asarray(w64.cpu() if hasattr(w64, "cpu") else w64)
        w16 = np.asarray(w16.cpu() if hasattr(w16, "cpu") else w16)
        rows.append({"defocus_nm": df_nm, "tcc_split16_ms": ms16, "tcc_fp64_ms": ms64,
                     "gram_split16_kernel_ms": gram_ms[-1],
                     "tcc_split16_kernels_ms": {k: v[1] for k, v in rep.items()},
                     "eig_top32_ms": mse, "eig_top32_householder_ms": mshh,
                     "tcc_rel_fro_err_split16_vs_fp64":
                         float(np.linalg.norm(b - a) / np.linalg.norm(a)),
                     "lambda_max_rel_err_split16": float(np.abs(w16 - w64).max() / w64[0]),
                     "lambda0": float(w64[0])})
    # the PSD path's A X products (n x n times n x 64, complex: 8 flops a MAC)
    zg = eig_kernels["subspace"]["k_zgemm"]
    zflops = 8.0 * T * T * 64
    out["eig_psd_products"] = zg["calls"]
    out["eig_psd_zgemm_tflops"] = zflops * zg["calls"] / (zg["ms"] / 1e3) / 1e12
    out["eig_psd_zgemm_frac_fp64_peak"] = out["eig_psd_zgemm_tflops"] / FP64_PEAK_TFLOPS
    out["eig_psd_product_tflops_whole_solve"] = (zflops * zg["calls"] /
                                                 (rows[2]["eig_top32_ms"] / 1e3) / 1e12)
    # the five FP64 TCCs in one batched solve (every launch covers the five)
    ctx.eig_topk_batch(t5, 32, psd=True)
    ms_batch = timer(lambda: ctx.eig_topk_batch(t5, 32, psd=True), 1)
    out["eig_top32_batch5_ms"] = ms_batch
    out["eig_top32_batch5_ms_per_matrix"] = ms_batch / t5.shape[0]
    # The reference accumulates the upper triangle (klitho_tcc.cpp:230-266):
    # 8 flops per complex MAC, T (T + 1) / 2 entries; the split issues 3 products.
    flops = 8.0 * used * T * (T + 1) / 2
    best = min(gram_ms)
    out.update({"rows": rows, "gram_flops_upper_triangle": flops,
                "gram_kernel_ms_best": best,
                "split16_tflops_algorithmic": flops / (best / 1e3) / 1e12,
                "split16_mfma_utilisation": 3 * flops / (best / 1e3) / 1e12 / MFMA_F16_PEAK_TFLOPS,
                "eig_top32_kernels_at_0nm": eig_kernels,
                "eig_note": "eig_top32_ms: the positive-semidefinite path (KLT_EIG_PSD, "
                            "block subspace iteration on A^2 with CholQR and Rayleigh-Ritz) "
                            "that the TCC tool and the source-mask loop use; "
                            "eig_top32_householder_ms: the general Hermitian path; "
                            "eig_psd_zgemm_tflops: the A X products (k_zgemm, FP64 MFMA) "
                            "over their own kernel time, against the FP64 peak",
                "note": "tcc_*_ms: whole klt_tcc_build calls (source set-up on the host, pupil, "
                        "operands, Gram) into device memory; gram_split16_kernel_ms: the Gram "
                        "and its split-K reduction kernels alone (HIP events)"})
    return out


def _dipole_tcc(ctx, torch, precision):
    """The T = 1225 blurred-dipole TCC (no defocus) on the device, and a
    function that rebuilds it."""
    src = blurred_dipole()
    sig = min(klt.outer_sigma(src), 1.0)
    Nx, Ny = klt.tcc_sizing(W.NA, W.PERIOD_35, W.PERIOD_35, sig)
    T = (2 * Nx + 1) * (2 * Ny + 1)
    dsrc = torch.from_numpy(src).cuda()
    t = torch.empty((T, T, 2), dtype=torch.float64, device="cuda")

    def build():
        ctx.tcc_build(dsrc, sig, W.NA, 0.0, W.PERIOD_35, W.PERIOD_35, Nx, Ny, precision, NMED,
                      out=t)
    build()
    return T, t, build


def section_tcc_split16(ctx, timer, torch, steps, warmup):
    """Split-FP16 builds of the T = 1225 blurred-dipole TCC, alone: `steps`
    timed builds after `warmup`."""
    T, _, build = _dipole_tcc(ctx, torch, 1)
    for _ in range(warmup):
        build()
    ms = timer(build, steps)
    log(f"tcc_split16 T={T}: {ms:.3f} ms per build")
    return {"section": "tcc_split16", "T": T, "builds": steps, "ms_per_build": ms}


def section_eig_psd(ctx, timer, torch, steps, warmup):
    """Top-32 FP64 eigensolve of the T = 1225 blurred-dipole TCC (no defocus) on
    the positive-semidefinite path, alone: `steps` timed solves after `warmup`."""
    T, t64, _ = _dipole_tcc(ctx, torch, 0)
    for _ in range(warmup):
        ctx.eig_topk(t64, 32, psd=True)
    ms = timer(lambda: ctx.eig_topk(t64, 32, psd=True), steps)
    log(f"eig_psd T={T}: {ms:.3f} ms per top-32 solve")
    return {"section": "eig_psd", "T": T, "k": 32, "solves": steps, "ms_per_solve": ms}


ADJ_KERNELS = ("k_band_mfma", "k_synth_resist", "k_synth_mfma")  # the adjoint's streams


def adjoint_batch(ctx, bank, timer, torch, dev, B):
    """klt_socs_backward on B relaxed 1024^2 masks (targets: the binary masks):
    ms per call (mean of 20 after 3 warm-up calls, the first of which tunes
    the shape) and per-kernel HIP-event times of one call."""
    tg = np.stack([W.manhattan(L, 3000 + i % 64) for i in range(B)])
    mk = np.stack([W.relaxed(tg[i], 4000 + i) for i in range(B)])
    dm = torch.from_numpy(mk).to(dev)
    dt = torch.from_numpy(tg).to(dev)
    for _ in range(3):
        klt.socs_backward(ctx, bank, dm, dt)
    # 20 calls: a 0.1 ms call is short enough that the host's issue of the
    # first one (after the synchronisation) would weigh on a few-call mean
    ms = timer(lambda: klt.socs_backward(ctx, bank, dm, dt), 20)
    # one call queued ahead of the profiled one, so that its first kernel does
    # not start on an idle device (the timer ends with a synchronisation)
    klt.socs_backward(ctx, bank, dm, dt)
    ctx.set_profiling(True)
    klt.socs_backward(ctx, bank, dm, dt)
    rep = ctx.profile_report()
    ctx.set_profiling(False)
    # the product with a given dL/dI (no forward passes), on the same masks
    dg = torch.randn_like(dm)
    for _ in range(3):
        klt.socs_vjp(ctx, bank, dm, dg)
    ms_vjp = timer(lambda: klt.socs_vjp(ctx, bank, dm, dg), 20)
    kern = {k: {"ms": v[1], "launches": v[0]} for k, v in rep.items()}
    # each streaming kernel moves 4 L^2 B algorithmic bytes: the mask read
    # (band), the target read (fused synthesis + resist band), the gradient
    # write (synthesis of the gradient)
    for k in ADJ_KERNELS:
        if k in kern:
            v = kern[k]
            v["algorithmic_bytes_per_launch"] = 4.0 * L * L * B
            v["achieved_gbs"] = 4.0 * L * L * B / (v["ms"] / v["launches"] / 1e3) / 1e9
    return dm, dt, {"batch": B, "mask_grads_per_s": B / (ms / 1e3), "ms_per_batch": ms,
                    "path_hbm_gbs": 12.0 * L * L * B / (ms / 1e3) / 1e9, "kernels": kern,
                    "vjp_mask_grads_per_s": B / (ms_vjp / 1e3), "vjp_ms_per_batch": ms_vjp}


ADJ_COUNTERS = os.path.join(ROOT, "profiles", "adjoint_1024_hbm_counters.json")


def adjoint_roofline(res):
    """Roofline object of the adjoint's dominant streaming kernel at one batch:
    its algorithmic bytes over its HIP-event duration, and (when a counter
    measurement of this kernel and batch is on file) its copy-calibrated
    FETCH + WRITE bytes per launch."""
    kern = res["kernels"]
    dom = max((k for k in ADJ_KERNELS if k in kern),
              key=lambda k: kern[k]["ms"] / kern[k]["launches"])
    v = kern[dom]
    traffic = None
    try:
        with open(ADJ_COUNTERS) as f:
            c = json.load(f)
        traffic = c["batches"][str(res["batch"])][dom]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        pass
    return {"kernel": dom, "batch": res["batch"], "bound": "hbm",
            "achieved": v["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": v["achieved_gbs"] / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": v["algorithmic_bytes_per_launch"]}


def bench_adjoint(ctx, bank, timer, torch, dev, src):
    """Config 4's batch (16 masks) with the source gradient, then 128 masks."""
    dm, dt, r16 = adjoint_batch(ctx, bank, timer, torch, dev, 16)
    ds = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
    # first source-gradient call outside the profiled one (buffers, first launches)
    klt.abbe_source_grad(ctx, dm, dt, ds, W.NA, 0.0, 14, 14, nmed=NMED)
    ctx.set_profiling(True)
    klt.abbe_source_grad(ctx, dm, dt, ds, W.NA, 0.0, 14, 14, nmed=NMED)
    kern_src = {k: v[1] for k, v in ctx.profile_report().items()}
    ctx.set_profiling(False)
    ms_src = timer(lambda: klt.abbe_source_grad(ctx, dm, dt, ds, W.NA, 0.0, 14, 14, nmed=NMED), 3)
    del dm, dt
    _, _, r128 = adjoint_batch(ctx, bank, timer, torch, dev, 128)
    out = dict(r16)
    out.update({"roofline": adjoint_roofline(r16), "source_grad_ms_per_batch": ms_src,
                "source_grad_kernels_ms": kern_src,
                "hbm_bytes_per_mask_grad": 12 * L * L,
                "batch128": dict(r128, roofline=adjoint_roofline(r128)),
                "note": "algorithmic bytes per mask gradient: mask + target read, gradient "
                        "written (12 MiB at 1024^2); the forward image is formed tile by tile "
                        "inside k_synth_resist and never stored"})
    return out


def section_adjoint(ctx, timer, torch, steps, warmup, batch):
    """klt_socs_backward alone on `batch` relaxed 1024^2 masks with the 24
    kernels of 29 x 29: `steps` timed calls after `warmup` (for profiling)."""
    _, kr, ki, sc = gpu_kernels(ctx)
    bank = ctx.kernels(kr, ki, sc, KSIZE, KSIZE)
    dev = torch.device("cuda:0")
    tg = np.stack([W.manhattan(L, 3000 + i % 64) for i in range(batch)])
    mk = np.stack([W.relaxed(tg[i], 4000 + i) for i in range(batch)])
    dm = torch.from_numpy(mk).to(dev)
    dt = torch.from_numpy(tg).to(dev)
    for _ in range(warmup):
        klt.socs_backward(ctx, bank, dm, dt)
    ms = timer(lambda: klt.socs_backward(ctx, bank, dm, dt), steps)
    log(f"adjoint batch {batch}: {ms:.3f} ms per call, {batch / (ms / 1e3):.0f} mask grads/s")
    bank.close()
    return {"section": "adjoint", "batch": batch, "calls": steps, "ms_per_call": ms,
            "mask_grads_per_s": batch / (ms / 1e3)}


WIDE_KERNELS = ("k_band_mfma", "k_synth_wide")  # the wide forward's streams


def _wide_row(ctx, timer, torch, dev, kr, ki, sc, N, side, B, B_adj):
    """Forward images/s over B seeded side^2 Manhattan masks and adjoint mask
    gradients/s over B_adj relaxed ones on the kernel half-size N > 15 path,
    with per-kernel HIP-event times of one forward call."""
    bank = ctx.kernels(kr, ki, sc, 2 * N + 1, 2 * N + 1)
    mh = np.stack([W.manhattan(side, 7000 + i) for i in range(min(B, 16))])
    masks = torch.from_numpy(mh).to(dev).repeat((B + len(mh) - 1) // len(mh), 1, 1)[:B]
    masks = masks.contiguous()
    images = torch.empty_like(masks)
    for _ in range(3):
        ctx.socs_forward(bank, masks, images)
    ms = timer(lambda: ctx.socs_forward(bank, masks, images), 10)
    ctx.socs_forward(bank, masks, images)  # (the profiled call does not start on an idle device)
    ctx.set_profiling(True)
    ctx.socs_forward(bank, masks, images)
    rep = ctx.profile_report()
    ctx.set_profiling(False)
    kern = {k: {"ms": v[1], "launches": v[0]} for k, v in rep.items()}
    for k in WIDE_KERNELS:
        if k in kern:
            kern[k]["achieved_gbs"] = 4.0 * side * side * B / (kern[k]["ms"] / 1e3) / 1e9
    dm, dt = masks[:B_adj].clone(), masks[:B_adj].clone()
    dm.add_(torch.rand_like(dm) * 0.2 - 0.1).clamp_(0.0, 1.0)
    for _ in range(3):
        klt.socs_backward(ctx, bank, dm, dt)
    ms_adj = timer(lambda: klt.socs_backward(ctx, bank, dm, dt), 10)
    bank.close()
    return {"N": N, "K": int(len(sc)), "side": side, "batch": B,
            "images_per_s": B / (ms / 1e3), "ms_per_batch": ms,
            "path_hbm_gbs": 8.0 * side * side * B / (ms / 1e3) / 1e9, "kernels": kern,
            "adjoint_batch": B_adj, "mask_grads_per_s": B_adj / (ms_adj / 1e3),
            "adjoint_ms_per_batch": ms_adj}


def bench_wide(ctx, timer, torch, dev):
    """Kernel half-sizes above 15 (the matrix-core wide path: 64-row band, 128-point
    coarse FFTs, K = 128 synthesis):
      sample  the reference's own test_tcc.sh imaging: 2048^2 period, annular 101
              grid 0.6 / 0.9, NA 0.8, defocus 0.2, 3 kernels of 33 x 33 (N = 16),
              built on the GPU; 32 masks of 2048^2 (adjoint: 8);
      n31     N = 31 (the largest half-size): 24 kernels of 63 x 63 from the annular
              51 grid at NA 1.35, period 2400 nm; 128 masks of 1024^2 (adjoint:
              128).
    (Their images and gradients are checked against the FP64 oracle in the
    tests.)"""
    out = {}
    src = klt.source_create(101, "Annular", 0.6, 0.9)
    sig = min(klt.outer_sigma(src), 1.0)
    Nx, Ny = klt.tcc_sizing(0.8, 2048, 2048, sig)
    tcc = ctx.tcc_build(src, sig, 0.8, 0.2, 2048, 2048, Nx, Ny, precision=0)
    kr, ki, sc = ctx.calc_kernels(tcc, 3)
    out["sample"] = _wide_row(ctx, timer, torch, dev, kr, ki, sc, Nx, 2048, 32, 8)
    src = annular_source()
    sig = min(klt.outer_sigma(src), 1.0)
    Nx, Ny = klt.tcc_sizing(W.NA, 2400, 2400, sig)
    tcc = ctx.tcc_build(src, sig, W.NA, 0.0, 2400, 2400, Nx, Ny, precision=0, nmed=NMED)
    kr, ki, sc = ctx.calc_kernels(tcc, 24)
    out["n31"] = _wide_row(ctx, timer, torch, dev, kr, ki, sc, Nx, 1024, 128, 128)
    return out


def bench_widen(ctx, timer, torch, dev, masks, images, src):
    """The rows either side of the imaging path, on the workload's own shapes:
      hopkins  klt_hopkins_image (calcImage + FT_c2r) of one 1024^2 mask from the
               FP64 TCC of the headline optics (T = 841);
      resist   klt_resist_levels over the headline batch's images (512 of
               1024^2: 4 bytes read and 1 written a pixel);
      smo      klt_smo_run on 16 relaxed 1024^2 Manhattan masks, 51^2 annular
               source, K = 24: each step builds the TCC, solves its top 24
               eigenpairs, images the masks, and forms the mask adjoint and the
               source gradient before the projected updates."""
    out = {}
    sig = min(klt.outer_sigma(src), 1.0)
    dsrc = torch.from_numpy(src).to(dev)
    tcc = ctx.tcc_build(dsrc, sig, W.NA, 0.0, W.PERIOD_29, W.PERIOD_29, 14, 14, precision=0,
                        nmed=NMED)
    m1 = masks[0]
    ctx.hopkins_image(m1, tcc, 14, 14)
    hms = timer(lambda: ctx.hopkins_image(m1, tcc, 14, 14), 5)
    out["hopkins"] = {"T": int(tcc.shape[0]), "side": int(m1.shape[-1]), "ms_per_image": hms}
    ctx.resist_levels(images)
    rms = timer(lambda: ctx.resist_levels(images), 5)
    nbytes = 5.0 * images.numel()
    out["resist_levels"] = {"images": int(images.shape[0]), "side": int(images.shape[-1]),
                            "ms": rms, "hbm_gbs": nbytes / (rms / 1e3) / 1e9}
    T = W.manhattan_batch(L, 16, 3000)
    m0 = np.stack([W.relaxed(T[i], 4000 + i) for i in range(16)])
    dm = torch.from_numpy(m0).to(dev)
    dt = torch.from_numpy(T).to(dev)
    S = klt.source_create(W.SRC_SIZE, "Annular", W.SIGMA_IN, W.SIGMA_OUT)
    S /= S.sum()
    klt.smo_run(ctx, dm, dt, S, W.NA, 1, nmed=NMED)  # (first call: tuning of its shapes)
    steps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist, _ = klt.smo_run(ctx, dm, dt, S, W.NA, steps, nmed=NMED)
    torch.cuda.synchronize()
    sms = (time.perf_counter() - t0) * 1e3 / steps
    out["smo"] = {"masks": 16, "side": L, "K": NK, "source": W.SRC_SIZE, "steps": steps,
                  "ms_per_step": sms, "loss_first": float(hist[0]), "loss_last": float(hist[-1])}
    return out


def copy_calibration(torch, timer, dev):
    """A torch device-to-device copy of 1 GiB (the runtime's copy kernel), whose
    known bytes calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev)
    bcopy = torch.empty_like(a)
    cms = timer(lambda: bcopy.copy_(a), 5)
    out = {"bytes_read": a.numel() * 4, "bytes_written": a.numel() * 4, "ms": cms,
           "gbs": 2 * a.numel() * 4 / (cms / 1e3) / 1e9}
    log(f"copy calibration: {out['gbs']:.0f} GB/s")
    return out


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_threads():
    """Worker threads for the all-core leg: the cores this process may use,
    at most 16 (one GPU's share of the host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_io_timing(O, img):
    """The host-side rows beside the reference's own code (oracle/_ref: K-Litho's
    source.cpp, mask.cpp and file_io.cpp compiled from its sources): a 1024^2
    whitespace-text mask read through createMask's Import, a generated
    LineSpace mask, the 101^2 annular source grid, image.bin written and read,
    the 24 kernel files and kernel_info.txt of klitho_tcc, and the three PNG
    views of one 1024^2 image; milliseconds, the median of 5 runs each."""
    import shutil
    import tempfile
    R = O.ref()
    out = {}
    d = tempfile.mkdtemp(prefix="klt_hostio_")

    def med(fn):
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            fn()
            ts.append(1e3 * (time.perf_counter() - t0))
        return float(np.median(ts))

    # (the Import path prints its file name, as the reference's tools do:
    # bench.py's stdout carries only the JSON line, so fd 1 is silenced here)
    sys.stdout.flush()
    saved, devnull = os.dup(1), os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        path = os.path.join(d, "mask.txt")
        np.savetxt(path, W.manhattan(L, 4242), fmt="%g")
        out["text_mask_import_ms"] = med(
            lambda: klt.mask_create(L, L, L, L, "Import", mask_file=path))
        if R is not None:
            buf = np.zeros(L * L, np.float32)
            out["text_mask_import_ms_reference"] = med(
                lambda: R.ref_create_mask(buf, L, L, L, L, b"Import", 1, 1, 0, path.encode(),
                                          1.0))
        # createMask's generated LineSpace mask, createSource's annular grid of
        # the TCC study (101^2), image.bin written and read back
        out["linespace_mask_ms"] = med(
            lambda: klt.mask_create(L, L, L, L, "LineSpace", 128, 128, 0))
        out["annular_source_101_ms"] = med(lambda: klt.source_create(101, "Annular", 0.6, 0.9))
        im = np.ascontiguousarray(img, np.float32)
        fb = os.path.join(d, "image.bin")
        out["image_bin_write_ms"] = med(lambda: klt.write_floats(fb, im))
        out["image_bin_read_ms"] = med(lambda: klt.read_floats(fb, L * L))
        if R is not None:
            out["linespace_mask_ms_reference"] = med(
                lambda: R.ref_create_mask(buf, L, L, L, L, b"LineSpace", 128, 128, 0, b"", 1.0))
            sbuf = np.zeros(101 * 101, np.float32)
            out["annular_source_101_ms_reference"] = med(
                lambda: R.ref_create_source(sbuf, 101, b"Annular", 0.6, 0.9, 0.1, 0.5, 1, 0.6,
                                            0.2, b""))
            out["image_bin_write_ms_reference"] = med(
                lambda: R.ref_write_float_bin(fb.encode(), im.ravel(), L * L))
            out["image_bin_read_ms_reference"] = med(
                lambda: R.ref_read_float_bin(fb.encode(), buf, L * L))
        # the kernel files klitho_tcc writes: 24 kernels of 29 x 29 (krn_*_r/i.bin)
        # and kernel_info.txt
        kk = np.random.default_rng(1).standard_normal((48, 29 * 29)).astype(np.float32)
        scl = np.linspace(1.0, 0.01, 24).astype(np.float32)

        def kernel_files(ref):
            for q in range(48):
                f = os.path.join(d, f"krn_{q // 2}_{'ri'[q % 2]}.bin")
                if ref:
                    R.ref_write_float_bin(f.encode(), kk[q], 29 * 29)
                else:
                    klt.write_floats(f, kk[q])
            f = os.path.join(d, "kernel_info.txt")
            if ref:
                R.ref_write_kernel_info(f.encode(), 29, 29, 24, scl)
            else:
                klt.write_kernel_info(f, 29, 29, scl)

        out["kernel_files_write_ms"] = med(lambda: kernel_files(False))
        if R is not None:
            out["kernel_files_write_ms_reference"] = med(lambda: kernel_files(True))
        for kind in ("gray", "7colors", "continuous"):
            p = os.path.join(d, f"{kind}.png")
            out[f"png_{kind}_ms"] = med(
                lambda: klt.write_png(kind, p, im, 0.0, klt.TARGET_INTENSITY))
            if R is not None:
                out[f"png_{kind}_ms_reference"] = med(
                    lambda: getattr(R, f"ref_png_{kind}")(p.encode(), L, L, im.ravel(), 0.0,
                                                          klt.TARGET_INTENSITY))
    finally:
        import ctypes
        ctypes.CDLL(None).fflush(None)
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
        shutil.rmtree(d, ignore_errors=True)
    return out


def cpu_baseline(ctx, bank, fused_img, mask, kr, ki, sc, budget_s):
    """The reference's CPU path on this host, timed on bounded samples:
    FT_r2c + calcSOCS + FI on FFTW3 (oracle/cpu_ref_path.cpp, the library
    loaded at run time) on 1 core and on all cores (one worker per core, each
    single threaded), calcTCC (the oracle's C restatement, 1 core) and
    calcKernels' zheevr_ top-32 at T = 1225 on 1 and all cores; plus the FP64
    oracle errors of the GPU image and mask gradient and the numpy port of the
    FFTW path.  `value` is the all-core images/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_ref as CR
    import oracle as O
    import ref_pipeline as RP
    N = (KSIZE - 1) // 2
    pool = np.stack([W.manhattan(L, 5000 + i) for i in range(16)])
    ref64 = O.socs_image(mask, kr, ki, sc, N, N)
    err = float(np.linalg.norm(fused_img - ref64) / np.linalg.norm(ref64))
    tg = W.manhattan(512, 77)
    mk = W.relaxed(tg, 78)
    _, g, _ = klt.socs_backward(ctx, bank, mk[None].copy(), tg[None].copy())
    _, _, g64 = O.socs_loss_grad(mk.astype(np.float64), tg, kr, ki, sc, N, N)
    gerr = float(np.linalg.norm(g[0] - g64) / np.linalg.norm(g64))
    nth = _host_threads()
    info = {"nproc": os.cpu_count(), "cpu_model": _cpu_model(), "threads_all_cores": nth,
            "library": CR.reason(), "host_io": host_io_timing(O, fused_img)}
    # numpy port of the FFTW procedure (secondary figure), ~1/4 of the budget
    t0, n = time.perf_counter(), 0
    while time.perf_counter() - t0 < 0.25 * budget_s:
        RP.socs_pipeline(pool[n % 16], kr, ki, sc, N, N)
        n += 1
    info["numpy_port_images_per_s_1core"] = n / (time.perf_counter() - t0)
    if not CR.available():
        cpu = {"value": info["numpy_port_images_per_s_1core"], "unit": "images/s", "cores": 1,
               "kind": "port", "sample": f"FFTW3/LAPACK {CR.reason()}; numpy pocketfft port of "
                                         f"FT_r2c+calcSOCS+FI, {n} seeded 1024^2 masks, 1 core",
               **info}
        return cpu, err, gerr
    one = CR.socs_images(mask, kr, ki, sc, N, N, threads=1)[0]
    info["image_rel_l2_vs_fp64_oracle"] = float(np.linalg.norm(one - ref64) / np.linalg.norm(ref64))
    # 1 core
    t0, n1 = time.perf_counter(), 0
    while time.perf_counter() - t0 < 0.25 * budget_s:
        CR.socs_images(pool[n1 % 16:n1 % 16 + 2], kr, ki, sc, N, N, threads=1)
        n1 += 2
    rate1 = n1 / (time.perf_counter() - t0)
    # all cores: one call of 2 masks per worker, repeated within the budget
    batch = np.concatenate([pool] * ((2 * nth + 15) // 16))[:2 * nth]
    t0, na = time.perf_counter(), 0
    while time.perf_counter() - t0 < 0.25 * budget_s:
        CR.socs_images(batch, kr, ki, sc, N, N, threads=nth)
        na += batch.shape[0]
    rate_all = na / (time.perf_counter() - t0)
    # calcTCC + zheevr_ top 32 at T = 1225 (blurred dipole, period 1280, 0 nm)
    src = blurred_dipole()
    sig = min(float(O.outer_sigma(src)), 1.0)
    t0 = time.perf_counter()
    t = O.tcc(src, sig, W.NA, 0.0, W.PERIOD_35, W.PERIOD_35, 17, 17, nmed=NMED)
    info["calc_tcc_T1225_ms_1core"] = 1e3 * (time.perf_counter() - t0)
    for th in (1, nth):
        t0 = time.perf_counter()
        CR.zheevr_topk(t, 32, threads=th)
        info[f"zheevr_top32_T1225_ms_{th}threads"] = 1e3 * (time.perf_counter() - t0)
    info["images_per_s_1core"] = rate1
    cpu = {"value": rate_all, "unit": "images/s", "cores": nth, "kind": "port",
           "sample": f"{na} (all cores) and {n1} (1 core) seeded 1024^2 Manhattan masks, K=24 "
                     "29x29: the reference's FT_r2c+calcSOCS+FI restated on FFTW3 "
                     "(FFTW_ESTIMATE, double) from the library below, one single-threaded "
                     "worker per core",
           **info}
    return cpu, err, gerr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline and plain path only")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--section", choices=["all", "eig_psd", "tcc_split16", "adjoint", "wide",
                                         "batch_uneven"],
                    default="all",
                    help="eig_psd: only --steps top-32 solves of the T = 1225 blurred-dipole "
                         "TCC on the positive-semidefinite path; tcc_split16: only --steps "
                         "split16 builds of that TCC (no defocus); adjoint: only --steps "
                         "klt_socs_backward calls on --batch 1024^2 masks; wide: only the "
                         "kernel half-size 16 and 31 rows of the wide path; batch_uneven: only "
                         "the 192-image mixed batch; their timing as a JSON line (for "
                         "profiling; no bench line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the SOCS "
              "batch shards trivially by image and is measured on one GPU's share.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = klt.Context(0)
    stream = torch.cuda.Stream()  # default priority; the library launches on it
    ctx.set_stream(stream.cuda_stream)
    torch.cuda.set_stream(stream)
    timer = Timer(torch, stream)
    if args.section == "adjoint":
        res = section_adjoint(ctx, timer, torch, args.steps, args.warmup, args.batch)
        if args.copy_calibration:
            res["copy_calibration"] = copy_calibration(torch, timer, dev)
        print(json.dumps(res))
        return
    if args.section == "wide":
        print(json.dumps(bench_wide(ctx, timer, torch, dev)))
        return
    if args.section == "batch_uneven":
        print(json.dumps(bench_batch_uneven(ctx, timer, torch, gpu_kernels(ctx)[0])))
        return
    if args.section != "all":
        fn = section_eig_psd if args.section == "eig_psd" else section_tcc_split16
        print(json.dumps(fn(ctx, timer, torch, args.steps, args.warmup)))
        return

    t0 = time.perf_counter()
    src, kr, ki, sc = gpu_kernels(ctx)
    log(f"kernels from the GPU TCC + eigensolver: lambda0 {sc[0]:.4f} "
        f"({time.perf_counter() - t0:.2f} s)")
    bank = ctx.kernels(kr, ki, sc, KSIZE, KSIZE)
    B = args.batch
    log(f"generating masks of {L}^2 (batch {B})")
    masks_h = W.manhattan_batch(L, min(B, 64), 1000)
    reps = (B + masks_h.shape[0] - 1) // masks_h.shape[0]
    masks = torch.from_numpy(masks_h).to(dev).repeat(reps, 1, 1)[:B].contiguous()
    images = torch.empty_like(masks)

    def step():
        ctx.socs_forward(bank, masks, images)

    for _ in range(args.warmup):
        step()
    w0 = time.perf_counter()
    ms = timer(step, args.steps)
    wall = time.perf_counter() - w0
    value = B / (ms / 1e3)
    log(f"batched fused: {ms:.3f} ms/step, {value:.1f} images/s (wall {wall:.2f} s)")
    fused_img0 = images[0].cpu().numpy()

    # per-kernel breakdown (HIP events bracketing each launch on the stream)
    ctx.set_profiling(True)
    prof_steps = 3
    for _ in range(prof_steps):
        step()
    rep = ctx.profile_report()
    ctx.set_profiling(False)
    kern = {k: {"ms_per_launch": v[1] / v[0], "launches": v[0] // prof_steps}
            for k, v in rep.items()}
    tot = sum(v[1] for v in rep.values()) / prof_steps
    for k in kern:
        kern[k]["share"] = kern[k]["ms_per_launch"] * kern[k]["launches"] / tot
    alg = {"k_band_mfma": 4.0 * L * L, "k_synth_mfma": 4.0 * L * L}  # bytes per image
    for k, v in kern.items():
        if k in alg:
            per_launch = alg[k] * B / v["launches"]
            v["algorithmic_bytes_per_launch"] = per_launch
            v["achieved_gbs"] = per_launch / (v["ms_per_launch"] / 1e3) / 1e9
    dom = max((k for k in kern if k in alg), key=lambda k: kern[k]["ms_per_launch"])
    ach = kern[dom]["achieved_gbs"]
    roof = {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, B // kern[dom]["launches"]),
            "algorithmic_bytes_per_launch": kern[dom]["algorithmic_bytes_per_launch"]}
    path_gbs = 8.0 * L * L * B / (ms / 1e3) / 1e9

    # what plain streaming kernels reach on the same buffers: a read-only
    # reduction of the masks, a fill and a copy of the images (2 GiB each)
    nbytes = 4.0 * masks.numel()
    rates = {}
    for name, fn in (("read", lambda: masks.sum()), ("write", lambda: images.fill_(0.5)),
                     ("copy", lambda: images.copy_(masks))):
        fn()  # first launch loads the kernel
        rates[name] = timer(fn, 5)
    r_ms, w_ms, c_ms = rates["read"], rates["write"], rates["copy"]
    stream_rates = {"read_gbs": nbytes / (r_ms / 1e3) / 1e9,
                    "write_gbs": nbytes / (w_ms / 1e3) / 1e9,
                    "copy_gbs": 2 * nbytes / (c_ms / 1e3) / 1e9,
                    "bytes": nbytes,
                    "note": "torch sum over the masks, fill_ and copy_ into the images"}
    log(f"streaming reference rates: read {stream_rates['read_gbs']:.0f}, write "
        f"{stream_rates['write_gbs']:.0f}, copy {stream_rates['copy_gbs']:.0f} GB/s")

    # plain single-instance path: one call per image
    one_m = masks[:1].clone()
    one_i = torch.empty_like(one_m)
    for _ in range(3):
        ctx.socs_forward(bank, one_m, one_i, plain=True)
    plain = 1.0 / (timer(lambda: ctx.socs_forward(bank, one_m, one_i, plain=True), 64) / 1e3)
    log(f"plain single-image path: {plain:.1f} images/s")

    extras = {}
    if not args.no_extras:
        adj = bench_adjoint(ctx, bank, timer, torch, dev, src)
        log(f"vjp (given dL/dI): {adj['vjp_mask_grads_per_s']:.1f} mask grads/s at 16, "
            f"{adj['batch128']['vjp_mask_grads_per_s']:.1f} at 128")
        log(f"adjoint: {adj['mask_grads_per_s']:.1f} mask grads/s at 16 "
            f"({adj['roofline']['kernel']} {adj['roofline']['frac']:.3f} of HBM peak), "
            f"{adj['batch128']['mask_grads_per_s']:.1f} at 128; source grad "
            f"{adj['source_grad_ms_per_batch']:.2f} ms / 16 masks")
        extras["adjoint"] = adj
        tce = bench_tcc_eig(ctx, timer, torch)
        log(f"tcc T={tce['T']}: split16 Gram {tce['gram_kernel_ms_best']:.3f} ms, "
            f"{tce['split16_tflops_algorithmic']:.1f} TF/s algorithmic, MFMA "
            f"{100 * tce['split16_mfma_utilisation']:.1f} %; eig top-32 "
            f"{tce['rows'][2]['eig_top32_ms']:.2f} ms (Householder path "
            f"{tce['rows'][2]['eig_top32_householder_ms']:.1f} ms), five in one batched call "
            f"{tce['eig_top32_batch5_ms']:.1f} ms; A X products {tce['eig_psd_zgemm_tflops']:.1f} "
            f"TF/s FP64 ({100 * tce['eig_psd_zgemm_frac_fp64_peak']:.0f} % of peak)")
        extras["tcc_eig"] = tce
        bu = bench_batch_uneven(ctx, timer, torch, src)
        log(f"uneven batch: {bu['images_per_s']:.0f} images/s ({bu['ms_per_batch']:.3f} ms per "
            f"192 images; grouped calls {bu['ms_per_batch_grouped_calls']:.3f} ms), p50 "
            f"{bu['p50_image_latency_us']:.0f} us, p99 {bu['p99_image_latency_us']:.0f} us")
        extras["batch_uneven"] = bu
        wd = bench_wide(ctx, timer, torch, dev)
        log(f"wide path: reference sample N={wd['sample']['N']} K={wd['sample']['K']} "
            f"2048^2: {wd['sample']['images_per_s']:.0f} images/s, "
            f"{wd['sample']['mask_grads_per_s']:.0f} mask grads/s; N={wd['n31']['N']} K=24 "
            f"1024^2: {wd['n31']['images_per_s']:.0f} images/s, "
            f"{wd['n31']['mask_grads_per_s']:.0f} mask grads/s")
        extras["wide"] = wd
        wn = bench_widen(ctx, timer, torch, dev, masks, images, src)
        log(f"Hopkins image {wn['hopkins']['ms_per_image']:.2f} ms (T={wn['hopkins']['T']}, "
            f"1024^2); resist levels {wn['resist_levels']['ms']:.3f} ms for 512 images "
            f"({wn['resist_levels']['hbm_gbs']:.0f} GB/s); SMO step "
            f"{wn['smo']['ms_per_step']:.2f} ms (16 masks of 1024^2, K=24)")
        extras["widen"] = wn
    if args.copy_calibration:
        extras["copy_calibration"] = copy_calibration(torch, timer, dev)

    cpu = err = gerr = None
    if not args.no_cpu:
        cpu, err, gerr = cpu_baseline(ctx, bank, fused_img0, masks_h[0], kr, ki, sc,
                                      args.cpu_seconds)
        hio = cpu["host_io"]
        log(f"host I/O: text mask import {hio['text_mask_import_ms']:.1f} ms (reference "
            f"{hio.get('text_mask_import_ms_reference', float('nan')):.1f}), PNG views "
            f"{hio['png_gray_ms']:.1f} / {hio['png_7colors_ms']:.1f} / "
            f"{hio['png_continuous_ms']:.1f} ms (reference "
            f"{hio.get('png_gray_ms_reference', float('nan')):.1f} / "
            f"{hio.get('png_7colors_ms_reference', float('nan')):.1f} / "
            f"{hio.get('png_continuous_ms_reference', float('nan')):.1f})")
        log(f"cpu baseline: {cpu['value']:.2f} images/s on {cpu['cores']} cores "
            f"({cpu.get('images_per_s_1core', 0):.1f} on 1); GPU image rel-L2 vs FP64 {err:.2e}, "
            f"mask gradient rel-L2 vs FP64 {gerr:.2e}")

    out = {
        "metric": "SOCS images/s at 1024^2 (K=24) + adjoint grads/s; rel-L2 err vs FP64 oracle",
        "value": value, "unit": "images/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "split-fp16/fp32",
        "data": "synthetic",
        "config": {"workload": "SOCS imaging, 1024x1024 seeded Manhattan masks, K=24 kernels "
                               "29x29 (annular 0.6/0.9, NA 1.35, period 1088 nm)",
                   "batch": B, "L": L, "K": NK, "kernel_size": KSIZE,
                   "precision": "split hi+lo FP16 matrix-core DFT passes, FP32 accumulation, "
                                "FP32 coarse FFTs",
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "image_rel_l2_vs_fp64": err,
        "grad_rel_l2_vs_fp64": gerr,
        "path_hbm_gbs": path_gbs,
        "plain_images_per_s": plain,
        "speedup_vs_plain": value / plain,
        "kernels": kern,
        "hbm_stream_rates": stream_rates,
    }
    out["tuner"] = [{"pass": r[0], "batch": r[1], "Lx": r[2], "Ly": r[3], "N": r[4], "K": r[6],
                     "band_rows": r[10], "band_column_splits": r[7], "synth_row_splits": r[8],
                     "kernel_groups": r[9], "ms": r[11]} for r in ctx.tune_report()]
    out.update(extras)
    print(json.dumps(out))
    bank.close()
    ctx.close()


if __name__ == "__main__":
    main()
