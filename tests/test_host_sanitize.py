"""The host side of the boundary under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): host_io.cpp, png_io.cpp and the two
tools' argument and file handling, built with g++ -fsanitize=address,undefined
(the device entry points replaced by stubs that report no GPU,
tests/sanitize/gpu_stub.cpp).  The reference itself writes out of bounds in
FT_c2r (K-Litho-TCC/klitho_tcc.cpp:162-168), which is why this runs at all.

* tests/sanitize/host_harness.cpp: sources of every type and tiny sizes,
  masks, imports of short / long / malformed / missing files, kernel_info with
  missing lines and a small capacity, the PNG writers on NaN / infinite / flat
  data;
* klitho_tcc / klitho_socs: wrong argument counts, invalid sizes, NA out of
  range, text and binary mask imports, the shipped import source and kernel
  set, a missing kernel directory, even kernel sizes; each run ends with the
  reference's message and exit status and no sanitizer report.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "k-litho_amd", "csrc")
HOST = os.path.join(ROOT, "k-litho_amd", "host")
SAN = os.path.join(ROOT, "tests", "sanitize")
GOLD = os.path.join(ROOT, "tests", "golden")
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
LIBSRC = [os.path.join(CSRC, f) for f in ("host_io.cpp", "png_io.cpp", "klt_error.cpp")]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=86",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=87")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("asan")
    out = {}
    for name, srcs in (("harness", [os.path.join(SAN, "host_harness.cpp")]),
                       ("klitho_tcc", [os.path.join(HOST, "klitho_tcc.cpp"),
                                       os.path.join(SAN, "gpu_stub.cpp")]),
                       ("klitho_socs", [os.path.join(HOST, "klitho_socs.cpp"),
                                        os.path.join(SAN, "gpu_stub.cpp")])):
        exe = str(d / name)
        r = subprocess.run(["g++", *FLAGS, "-o", exe, *srcs, *LIBSRC], capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        out[name] = exe
    return out


def _clean(r):
    text = r.stdout + r.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text, text
    assert "LeakSanitizer" not in text, text
    assert r.returncode not in (86, 87)


def _run(exe, args, cwd):
    return subprocess.run([exe, *map(str, args)], cwd=cwd, capture_output=True, text=True,
                          env=ENV, timeout=300)


@pytest.fixture
def wd(tmp_path):
    os.makedirs(tmp_path / "out" / "kernels" / "png")
    return tmp_path


def test_host_harness_clean(bins, wd):
    r = _run(bins["harness"], [str(wd)], wd)
    _clean(r)
    assert r.returncode == 0, r.stdout + r.stderr


def _tcc(L=128, ms=96, mtype="LineSpace", mfile="-", n=51, stype="Annular", sfile="-", NA=0.9):
    return [L, L, ms, ms, mtype, 8, 8, 0, mfile, n, stype, 0.6, 0.9, 0.2, 0.7, 1, 0.6, 0.2, sfile,
            NA, 0.1, 8]


def test_klitho_tcc_argument_and_file_paths_clean(bins, wd):
    np.savetxt(wd / "m.txt", (np.arange(96 * 96) % 3 == 0).astype(np.float32).reshape(96, 96))
    np.ones(96 * 90, np.float32).tofile(wd / "short.bin")  # short: zero-padded with a warning
    cases = [
        ([1, 2, 3], "Expected 23 arguments."),
        (_tcc(L=64, ms=96), "Invalid input parameters"),
        (_tcc(NA=1.5), "NA must be in the range"),
        (_tcc(), "no HIP device visible"),
        (_tcc(mtype="Import", mfile="m.txt"), "no HIP device visible"),
        (_tcc(mtype="Import", mfile="short.bin"), "no HIP device visible"),
        (_tcc(mtype="Import", mfile="missing.bin"), "Error"),
        (_tcc(n=101, stype="Import", sfile=os.path.join(GOLD, "src_test1_size101.bin")),
         "no HIP device visible"),
        (_tcc(n=3, stype="Point"), "no HIP device visible"),
    ]
    for args, msg in cases:
        r = _run(bins["klitho_tcc"], args, wd)
        _clean(r)
        assert r.returncode == 1 and msg in r.stdout + r.stderr, (args, r.stdout + r.stderr)
    assert os.path.exists(wd / "out" / "source.png") and os.path.exists(wd / "out" / "mask.png")


def test_klitho_socs_argument_and_file_paths_clean(bins, wd):
    kd = wd / "k"
    os.makedirs(kd)
    for f in os.listdir(GOLD):
        if f.startswith("krn_") or f == "kernel_info.txt":
            shutil.copy(os.path.join(GOLD, f), kd / f)
    ev = wd / "even"
    os.makedirs(ev)
    with open(ev / "kernel_info.txt", "w") as f:
        f.write("- Kernel Size: 4x5\n- Number of Kernels: 1\nEigenvalue 0: 1\n")
    np.savetxt(wd / "m.txt", np.ones((70, 70), np.float32))
    cases = [
        ([1, 2], "Expected 11 arguments."),
        ([64, 64, 80, 80, "LineSpace", 4, 4, 0, "-", "./k/"], "Invalid input parameters"),
        ([80, 80, 70, 70, "LineSpace", 4, 4, 0, "-", "./nokernels/"], "Error opening file"),
        ([80, 80, 70, 70, "LineSpace", 4, 4, 0, "-", "./even/"], "Kernel size must be odd."),
        ([80, 80, 70, 70, "Import", 0, 0, 0, "m.txt", "./k/"], "no HIP device visible"),
        ([60, 60, 60, 60, "LineSpace", 4, 4, 0, "-", "./k/"], "ERROR in FI()"),
    ]
    for args, msg in cases:
        r = _run(bins["klitho_socs"], args, wd)
        _clean(r)
        assert r.returncode == 1 and msg in r.stdout + r.stderr, (args, r.stdout + r.stderr)
