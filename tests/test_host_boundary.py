"""Host side of the drop-in boundary (libklt.so, CPU only).

Sources, masks, sizing and the file formats must be bit-identical to the
reference: they are compared with K-Litho's own source.cpp / mask.cpp /
file_io.cpp compiled into oracle/_ref (skipped when that was not built) and
with the oracle's restatement.
"""
import ctypes as C
import os

import numpy as np
import pytest

import klt
import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")

SOURCES = [
    ("Annular", 51, dict(inner=0.6, outer=0.9)),
    ("Annular", 101, dict(inner=0.4, outer=0.6)),
    ("Dipole", 101, dict(radius=0.2, offset=0.7, on_x_axis=1)),
    ("Dipole", 77, dict(radius=0.15, offset=0.5, on_x_axis=0)),
    ("CrossQuadrupole", 101, dict(radius=0.1, offset=0.5)),
    ("Point", 101, dict(pt_x=0.6, pt_y=0.2)),
    ("Point", 51, dict(pt_x=-0.5, pt_y=0.7)),
    ("Bogus", 21, dict()),
]


@pytest.mark.parametrize("kind,n,kw", SOURCES)
def test_source_create_matches_reference(kind, n, kw):
    mine = klt.source_create(n, kind, **kw)
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    ref = np.zeros(n * n, np.float32)
    R.ref_create_source(ref, n, kind.encode(), kw.get("inner", 0.6), kw.get("outer", 0.9),
                        kw.get("radius", 0.1), kw.get("offset", 0.5), kw.get("on_x_axis", 1),
                        kw.get("pt_x", 0.6), kw.get("pt_y", 0.2), b"")
    assert np.array_equal(mine.ravel().view(np.uint32), ref.view(np.uint32))


def test_source_import_bin_and_txt(tmp_path):
    p = os.path.join(GOLD, "src_test1_size101.bin")
    a = klt.source_create(101, "Import", source_file=p)
    assert abs(float(a.astype(np.float64).sum()) - 1.0) < 1e-4 and int((a > 0).sum()) == 1576
    txt = tmp_path / "src.txt"
    raw = np.fromfile(p, np.float32)
    txt.write_text("\n".join(" ".join(repr(float(v)) for v in raw[r * 101:(r + 1) * 101])
                             for r in range(101)))
    b = klt.source_create(101, "Import", source_file=str(txt))
    assert np.array_equal(a, b)
    R = O.ref()
    if R is not None:
        ref = np.zeros(101 * 101, np.float32)
        R.ref_create_source(ref, 101, b"Import", 0, 0, 0, 0, 0, 0, 0, str(txt).encode())
        assert np.array_equal(b.ravel(), ref)


def test_source_missing_file_is_an_io_error():
    with pytest.raises(klt.KltError) as e:
        klt.source_create(11, "Import", source_file="/nonexistent/src.bin")
    assert e.value.code == -5 and "Failed to open" in e.value.msg


@pytest.mark.parametrize("L,ms,lw,sw,hz,dose", [(512, 512, 32, 32, 0, 1.0),
                                                (256, 200, 10, 6, 1, 1.0),
                                                (65, 33, 3, 5, 0, 0.7)])
def test_linespace_mask_matches_reference(L, ms, lw, sw, hz, dose):
    mine = klt.mask_create(L, L, ms, ms, "LineSpace", lw, sw, hz, "", dose)
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built")
    ref = np.zeros(L * L, np.float32)
    R.ref_create_mask(ref, L, L, ms, ms, b"LineSpace", lw, sw, hz, b"", dose)
    assert np.array_equal(mine.ravel(), ref)


def test_import_mask_with_size_mismatch(tmp_path):
    rng = np.random.default_rng(0)
    m = rng.random((40, 30)).astype(np.float32)
    p = tmp_path / "m.bin"
    m.tofile(p)
    full = klt.mask_create(64, 50, 30, 40, "Import", mask_file=str(p))
    assert np.array_equal(full[5:45, 17:47], m)
    assert full.sum() == pytest.approx(m.sum(), rel=1e-6)
    # file shorter than the declared mask: zero padded, as the reference does
    short = klt.mask_create(64, 64, 40, 40, "Import", mask_file=str(p))
    inner = np.zeros(1600, np.float32)
    inner[:1200] = m.ravel()
    assert np.array_equal(short[12:52, 12:52], inner.reshape(40, 40))
    R = O.ref()
    if R is not None:
        ref = np.zeros(64 * 64, np.float32)
        R.ref_create_mask(ref, 64, 64, 40, 40, b"Import", 1, 1, 0, str(p).encode(), 1.0)
        assert np.array_equal(short.ravel(), ref)


def test_outer_sigma_and_sizing_match_oracle():
    for kind, n, kw in SOURCES[:6]:
        s = klt.source_create(n, kind, **kw)
        assert klt.outer_sigma(s) == O.outer_sigma(s)
        for NA, L in [(1.35, 1088), (0.8, 2048), (0.95, 700), (1.35, 1280)]:
            assert klt.tcc_sizing(NA, L, L, klt.outer_sigma(s)) == O.sizing(
                NA, L, L, klt.outer_sigma(s))


def test_kernel_info_round_trip_and_reference_reader(tmp_path):
    sc = np.array([78.79903, 37.28234, 37.28229, 0.0012345678], np.float32)
    p = str(tmp_path / "kernel_info.txt")
    klt.write_kernel_info(p, 29, 29, sc)
    txt = open(p).read()
    assert txt.splitlines()[:5] == ["# Kernel Information", "## Kernel Size and Count",
                                    "- Kernel Size: 29x29", "- Number of Kernels: 4",
                                    "## Eigenvalues of Each Kernel"]
    assert "Eigenvalue 0: 78.799\n" in txt and "Eigenvalue 3: 0.00123457\n" in txt
    sx, sy, got = klt.read_kernel_info(p)
    assert (sx, sy) == (29, 29) and np.allclose(got, sc, rtol=1e-5)
    R = O.ref()
    if R is not None:
        q = str(tmp_path / "ref_info.txt")
        R.ref_write_kernel_info(q.encode(), 29, 29, 4, sc)
        assert open(q).read() == txt
        a, b = C.c_int(), C.c_int()
        buf = np.zeros(8, np.float32)
        assert R.ref_read_kernel_info(p.encode(), C.byref(a), C.byref(b), buf, 8) == 4
        assert np.array_equal(buf[:4], got)


def test_shipped_kernel_info_fixture():
    sx, sy, sc = klt.read_kernel_info(os.path.join(GOLD, "kernel_info.txt"))
    assert (sx, sy, len(sc)) == (33, 33, 3)
    assert list(sc) == [np.float32(78.9658), np.float32(37.2823), np.float32(37.2823)]


def test_float_file_round_trip(tmp_path):
    d = np.random.default_rng(1).random(1000).astype(np.float32)
    p = str(tmp_path / "x.bin")
    klt.write_floats(p, d)
    assert np.array_equal(np.fromfile(p, np.float32), d)
    assert np.array_equal(klt.read_floats(p, 1000), d)
    assert np.array_equal(klt.read_floats(p, 1200)[:1000], d)
    with pytest.raises(klt.KltError):
        klt.write_floats("/nonexistent/dir/x.bin", d)
