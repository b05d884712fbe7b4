"""klitho_tcc / klitho_socs tools: argument handling, exit codes and messages
(CPU), and the end-to-end file pipeline on the GPU
(klitho_tcc -> out/kernels/ -> klitho_socs -> out/image.bin) against the
FP64 oracle: Hopkins image.bin of klitho_tcc, kernel_info.txt eigenvalues,
kernel unit norms, and the SOCS image.bin of klitho_socs.
"""
import os
import subprocess

import numpy as np
import pytest

import klt
import oracle as O
from conftest import relerr

BIN = os.path.join(klt.PKG, "bin")
TCC = os.path.join(BIN, "klitho_tcc")
SOCS = os.path.join(BIN, "klitho_socs")


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    return subprocess.run([exe, *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          env=e, timeout=300)


def _tcc_args(L=512, ms=512, mtype="LineSpace", lw=32, sw=32, hz=0, mfile="-", n=51,
              stype="Annular", inner=0.6, outer=0.9, radius=0.1, offset=0.5, onx=1, px=0.6,
              py=0.2, sfile="-", NA=0.9, df=0.0, nk=24):
    return [L, L, ms, ms, mtype, lw, sw, hz, mfile, n, stype, inner, outer, radius, offset, onx,
            px, py, sfile, NA, df, nk]


@pytest.fixture
def workdir(tmp_path):
    os.makedirs(tmp_path / "out" / "kernels" / "png")
    return tmp_path


def test_tcc_argument_count(workdir):
    r = _run(TCC, [1, 2, 3], workdir)
    assert r.returncode == 1
    assert "Invalid number of arguments: 4. Expected 23 arguments." in r.stderr
    assert "Lithography Simulation initiated" in r.stdout


def test_tcc_invalid_sizes_and_na(workdir):
    r = _run(TCC, _tcc_args(L=256, ms=512), workdir)
    assert r.returncode == 1 and "Invalid input parameters" in r.stderr
    r = _run(TCC, _tcc_args(NA=1.35), workdir)
    assert r.returncode == 1 and "NA must be in the range [0, 1]." in r.stderr
    r = _run(TCC, _tcc_args(NA=1.5), workdir, {"KLT_IMMERSION_INDEX": "1.44"})
    assert r.returncode == 1 and "NA must be in the range [0, 1.44]." in r.stderr


def test_tcc_non_numeric_argument_aborts(workdir):
    args = _tcc_args()
    args[0] = "abc"
    r = _run(TCC, args, workdir)
    assert r.returncode != 0  # std::stoi throws, uncaught: abort, as in the reference


def test_socs_argument_errors(workdir):
    r = _run(SOCS, [1, 2], workdir)
    assert r.returncode == 1 and "Expected 11 arguments." in r.stderr
    r = _run(SOCS, [64, 64, 80, 80, "LineSpace", 4, 4, 0, "-", "./k/"], workdir)
    assert r.returncode == 1 and "Invalid input parameters" in r.stderr
    r = _run(SOCS, [64, 64, 64, 64, "LineSpace", 4, 4, 0, "-", "./nokernels/"], workdir)
    assert r.returncode == 1 and "Error opening file for reading" in r.stderr
    kd = workdir / "even"
    os.makedirs(kd)
    klt.write_kernel_info(str(kd / "kernel_info.txt"), 4, 5, np.ones(1, np.float32))
    r = _run(SOCS, [64, 64, 64, 64, "LineSpace", 4, 4, 0, "-", str(kd) + "/"], workdir)
    assert r.returncode == 1 and "Kernel size must be odd." in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("NA,df,stype", [(0.9, 0.0, "Annular"), (0.8, 0.2, "Dipole")])
def test_tcc_then_socs_end_to_end(workdir, NA, df, stype):
    print(f"case cli NA={NA} df={df} {stype}", flush=True)
    L = 512
    kw = dict(L=L, ms=400, NA=NA, df=df, stype=stype, radius=0.2, offset=0.7, nk=12)
    r = _run(TCC, _tcc_args(**kw), workdir)
    assert r.returncode == 0, r.stderr + r.stdout
    # oracle: same source, mask, TCC, Hopkins image, eigenvalues
    if stype == "Annular":
        src = O.source("Annular", 51, inner=0.6, outer=0.9)
    else:
        src = O.source("Dipole", 51, radius=0.2, offset=0.7, onx=1)
    sig = min(O.outer_sigma(src), 1.0)
    Nx, Ny = O.sizing(NA, L, L, sig)
    assert f"Nx: {Nx}, Ny: {Ny}" in r.stdout
    inner = np.zeros(400 * 400, np.float32)
    O.lib().or_mask_linespace(inner, 400, 400, 32, 32, 0)
    mask = np.zeros(L * L, np.float32)
    O.lib().or_mask_place(inner, 400, 400, mask, L, L, 1.0)
    mask = mask.reshape(L, L)
    t = O.tcc(src, sig, NA, df, L, L, Nx, Ny)
    img = np.fromfile(workdir / "out" / "image.bin", np.float32).reshape(L, L)
    hop = O.hopkins_image(mask, t, Nx, Ny)
    assert relerr(img, hop) < 1e-5
    sx, sy, sc = klt.read_kernel_info(str(workdir / "out" / "kernels" / "kernel_info.txt"))
    assert (sx, sy) == (2 * Nx + 1, 2 * Ny + 1) and len(sc) == 12
    wo, vo = O.eigh_topk(t, 12)
    assert np.abs(sc - wo).max() < 1e-5 * wo[0]  # 6 printed digits
    T = sx * sy
    kr = np.stack([np.fromfile(workdir / "out" / "kernels" / f"krn_{i}_r.bin", np.float32)
                   for i in range(12)])
    ki = np.stack([np.fromfile(workdir / "out" / "kernels" / f"krn_{i}_i.bin", np.float32)
                   for i in range(12)])
    assert kr.shape == (12, T)
    assert np.allclose((kr.astype(np.float64) ** 2 + ki.astype(np.float64) ** 2).sum(1), 1,
                       atol=1e-5)
    # the reference's PNG views: sizes, and image_s.png = the 7-level classes
    # of image.bin against TARGET_INTENSITY (pinned to the reference's writer
    # in test_png_resist.py)
    from png_decode import read_png
    seven = np.array([(0, 0, 255), (0, 127, 255), (0, 255, 127), (0, 255, 0), (127, 255, 0),
                      (255, 127, 0), (255, 0, 0)], np.uint8)
    assert read_png(workdir / "out" / "image_s.png").shape == (L, L, 3)
    assert np.array_equal(read_png(workdir / "out" / "image_s.png"),
                          seven[O.resist_levels(img)])
    assert read_png(workdir / "out" / "image.png").shape == (L, L, 3)
    assert read_png(workdir / "out" / "mask.png").shape == (L, L)
    assert read_png(workdir / "out" / "source.png").shape == (51, 51)
    assert read_png(workdir / "out" / "tcc_r.png").shape == (T, T, 3)
    for i in range(12):
        assert read_png(workdir / "out" / "kernels" / "png" / f"krn_{i}_i.png").shape == (sy, sx, 3)
    # klitho_socs on the written kernels (both paths)
    for path in ("fused", "plain"):
        os.remove(workdir / "out" / "image.bin")
        r = _run(SOCS, [L, L, 400, 400, "LineSpace", 32, 32, 0, "-", "./out/kernels/"], workdir,
                 {"KLT_SOCS_PATH": path})
        assert r.returncode == 0, r.stderr + r.stdout
        simg = np.fromfile(workdir / "out" / "image.bin", np.float32).reshape(L, L)
        assert relerr(simg, O.socs_image(mask, kr, ki, sc, Nx, Ny)) < 1e-5


@pytest.mark.gpu
def test_reference_sample_workflow(workdir):
    """The reference's own sample (K-Litho-TCC/test_tcc.sh:4-49): 2048^2
    period, 1024^2 line/space mask 128/128 (the script's LineSpace values; its
    Import file is not shipped), annular 101 grid 0.6 / 0.9, NA 0.8, defocus
    0.2, 3 kernels: N = 16, 33 x 33.  klitho_tcc's Hopkins image and
    eigenvalues against the oracle's, then klitho_socs on the written kernels
    (the matrix-core path for N = 16, and the plain path) against the oracle's
    SOCS image of the same kernels."""
    print("case cli reference sample workflow 2048^2 N=16", flush=True)
    L, ms = 2048, 1024
    r = _run(TCC, _tcc_args(L=L, ms=ms, lw=128, sw=128, n=101, NA=0.8, df=0.2, nk=3), workdir)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Nx: 16, Ny: 16" in r.stdout
    src = O.source("Annular", 101, inner=0.6, outer=0.9)
    sig = min(O.outer_sigma(src), 1.0)
    inner = np.zeros(ms * ms, np.float32)
    O.lib().or_mask_linespace(inner, ms, ms, 128, 128, 0)
    mask = np.zeros(L * L, np.float32)
    O.lib().or_mask_place(inner, ms, ms, mask, L, L, 1.0)
    mask = mask.reshape(L, L)
    t = O.tcc(src, sig, 0.8, 0.2, L, L, 16, 16)
    img = np.fromfile(workdir / "out" / "image.bin", np.float32).reshape(L, L)
    assert relerr(img, O.hopkins_image(mask, t, 16, 16)) < 1e-5
    sx, sy, sc = klt.read_kernel_info(str(workdir / "out" / "kernels" / "kernel_info.txt"))
    assert (sx, sy, len(sc)) == (33, 33, 3)
    wo, _ = O.eigh_topk(t, 3)
    assert np.abs(sc - wo).max() < 1e-5 * wo[0]  # 6 printed digits
    kr = np.stack([np.fromfile(workdir / "out" / "kernels" / f"krn_{i}_r.bin", np.float32)
                   for i in range(3)])
    ki = np.stack([np.fromfile(workdir / "out" / "kernels" / f"krn_{i}_i.bin", np.float32)
                   for i in range(3)])
    ref = O.socs_image(mask, kr, ki, sc, 16, 16)
    for path in ("fused", "plain"):
        os.remove(workdir / "out" / "image.bin")
        r = _run(SOCS, [L, L, ms, ms, "LineSpace", 128, 128, 0, "-", "./out/kernels/"], workdir,
                 {"KLT_SOCS_PATH": path})
        assert r.returncode == 0, r.stderr + r.stdout
        simg = np.fromfile(workdir / "out" / "image.bin", np.float32).reshape(L, L)
        assert relerr(simg, ref) < 1e-5, path


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["default", "plain"])
def test_socs_reference_kernels_text_mask(workdir, path):
    """klitho_socs on the reference's own shipped kernel set (33 x 33, three
    kernels: half-size 16, the size its test_tcc.sh sample gives at period
    2048) with a 1024^2 Manhattan mask read from a whitespace text file
    (Import, common/mask.cpp:45-51 -> file_io.cpp:204-230), centred in a
    2048^2 period; image.bin against the oracle's SOCS image."""
    import shutil
    import workloads as W
    print(f"case cli reference kernels 33x33, 2048^2, text mask, path {path}", flush=True)
    kd = workdir / "ref_kernels"
    os.makedirs(kd)
    gold = os.path.join(os.path.dirname(__file__), "golden")
    for f in os.listdir(gold):
        if f.startswith("krn_") or f == "kernel_info.txt":
            shutil.copy(os.path.join(gold, f), kd / f)
    inner = W.manhattan(1024, 4242)
    np.savetxt(workdir / "mask.txt", inner, fmt="%g")
    env = {} if path == "default" else {"KLT_SOCS_PATH": "plain"}
    r = _run(SOCS, [2048, 2048, 1024, 1024, "Import", 0, 0, 0, "mask.txt", str(kd) + "/"],
             workdir, env)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Nx: 16,  Ny: 16" in r.stdout
    sx, sy, sc = klt.read_kernel_info(str(kd / "kernel_info.txt"))
    assert (sx, sy, len(sc)) == (33, 33, 3)
    kr = np.stack([np.fromfile(kd / f"krn_{i}_r.bin", np.float32) for i in range(3)])
    ki = np.stack([np.fromfile(kd / f"krn_{i}_i.bin", np.float32) for i in range(3)])
    mask = np.zeros(2048 * 2048, np.float32)
    O.lib().or_mask_place(np.ascontiguousarray(inner.ravel()), 1024, 1024, mask, 2048, 2048,
                          1.0)
    mask = mask.reshape(2048, 2048)
    img = np.fromfile(workdir / "out" / "image.bin", np.float32).reshape(2048, 2048)
    assert relerr(img, O.socs_image(mask, kr, ki, sc, 16, 16)) < 1e-5
