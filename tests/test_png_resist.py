"""PNG views and the resist step against the reference's own writers
(common/file_io.cpp:7-135, compiled into oracle/_ref with its vendored PNG
encoder): every pixel of klt_write_png_{gray,7colors,continuous} equals the
reference's PNG of the same data, and the oracle's 7-level classes equal the
colours of the reference's image_s.png view.  The data hold values on the
class boundaries k/6 of the range, below and above the range, NaN and
infinities, and the min == max case."""
import os

import numpy as np
import pytest

import klt
import oracle as O
from png_decode import read_png

SEVEN = {(0, 0, 255): 0, (0, 127, 255): 1, (0, 255, 127): 2, (0, 255, 0): 3,
         (127, 255, 0): 4, (255, 127, 0): 5, (255, 0, 0): 6}


def _data(rng, sx, sy, lo, hi):
    d = rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), (sy, sx)).astype(np.float32)
    flat = d.ravel()
    edges = np.float32(lo) + (np.arange(7, dtype=np.float32) / np.float32(6)) * np.float32(hi - lo)
    flat[:7] = edges
    flat[7:14] = np.nextafter(edges, np.float32(-np.inf))
    flat[14:17] = [np.nan, np.inf, -np.inf]
    return d


CASES = [(37, 23, 0.0, 0.225), (64, 64, -1.5, 2.0), (5, 9, 3.0, 3.0)]


@pytest.mark.parametrize("sx,sy,lo,hi", CASES)
@pytest.mark.parametrize("kind", ["gray", "7colors", "continuous"])
def test_png_views_equal_the_reference(tmp_path, sx, sy, lo, hi, kind):
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    d = _data(np.random.default_rng(sx * sy), sx, sy, lo, hi)
    mine, ref = str(tmp_path / "mine.png"), str(tmp_path / "ref.png")
    klt.write_png(kind, mine, d, lo, hi)
    getattr(R, f"ref_png_{kind}")(ref.encode(), sx, sy, np.ascontiguousarray(d.ravel()), lo, hi)
    a, b = read_png(mine), read_png(ref)
    assert a.shape == b.shape
    assert np.array_equal(a, b)


@pytest.mark.parametrize("sx,sy,lo,hi", CASES)
def test_oracle_levels_equal_reference_seven_colour_view(tmp_path, sx, sy, lo, hi):
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    d = _data(np.random.default_rng(7 + sx), sx, sy, lo, hi)
    path = str(tmp_path / "s.png")
    R.ref_png_7colors(path.encode(), sx, sy, np.ascontiguousarray(d.ravel()), lo, hi)
    rgb = read_png(path)
    lev = np.array([[SEVEN[tuple(int(c) for c in px)] for px in row] for row in rgb], np.uint8)
    assert np.array_equal(O.resist_levels(d, lo, hi), lev)
    if lo != hi:
        assert set(np.unique(lev)) == set(range(7))


def test_png_write_error(tmp_path):
    with pytest.raises(klt.KltError) as e:
        klt.write_png("gray", str(tmp_path / "no" / "dir" / "x.png"), np.zeros((2, 2), np.float32),
                      0.0, 1.0)
    assert "Failed to write PNG" in e.value.msg
