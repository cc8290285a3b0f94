"""The oracle re-runs the reference's own known-answer tests for the hot path.

Each test names the reference test it restates (file:line).  These pin the
oracle before the HIP kernels are compared with it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ---- crates/color/src/convert.rs -------------------------------------------

def test_div255_exhaustive_within_one():
    # convert.rs:316-326 div255_exhaustive
    L = O.lib()
    got = np.array([L.rro_div255(x) for x in range(0, 65536, 1)], np.int64)
    exp = np.minimum(np.round(np.arange(65536) / 255.0), 255)
    assert np.abs(got - exp).max() <= 1


def test_div255_products_and_identity():
    # convert.rs:329-341 div255_boundary_products, plus the identity the AA
    # composite relies on: div255(255*d) == d (shape 0 leaves the pixel alone).
    L = O.lib()
    for a in range(0, 256, 5):
        for b in range(256):
            assert abs(L.rro_div255(a * b) - round(a * b / 255.0)) <= 1
    for d in range(256):
        assert L.rro_div255(255 * d) == d
    assert L.rro_div255(65025) == 255 and L.rro_div255(0) == 0
    assert L.rro_div255(65535) == 255  # saturates


@pytest.mark.parametrize("x,fl,ce,ro", [
    (0.0, 0, 0, 0), (0.5, 0, 1, 1), (0.9, 0, 1, 1), (1.0, 1, 1, 1),
    (-0.1, -1, 0, 0), (-0.5, -1, 0, 0), (-0.6, -1, 0, -1)])
def test_splash_floor_ceil_round(x, fl, ce, ro):
    # convert.rs:377-392
    L = O.lib()
    assert (L.rro_splash_floor(x), L.rro_splash_ceil(x), L.rro_splash_round(x)) == (fl, ce, ro)


def test_splash_round_half_and_non_finite():
    # convert.rs:397-414
    L = O.lib()
    assert [L.rro_splash_round(v) for v in (0.5, -0.5, 1.5, -1.5)] == [1, 0, 2, -1]
    assert L.rro_splash_floor(math.inf) == 2**31 - 1
    assert L.rro_splash_floor(-math.inf) == -2**31
    assert L.rro_splash_floor(math.nan) == -2**31
    assert L.rro_splash_floor(1e300) == 2**31 - 1


# ---- crates/raster/src/fill/mod.rs -----------------------------------------

def test_aa_gamma_formula():
    # fill/mod.rs:516-538 aa_gamma_table_correct
    g = O.aa_gamma()
    exp = [int(math.floor((i / 16.0) ** 1.5 * 255.0 + 0.5)) for i in range(17)]
    assert list(g) == exp


# ---- crates/raster/src/xpath.rs tests (:506-680) -----------------------------

def seg(x0, y0, x1, y1):
    return O.add_segments(np.array([[x0, y0, x1, y1]], float))[0]


def test_add_segment_flags_and_slopes():
    s = seg(0, 5, 10, 5)  # horizontal_segment_not_flipped
    assert s.flags & O.SEG_HORIZ and not s.flags & O.SEG_FLIPPED and s.y0 == 5 and s.y1 == 5
    s = seg(0, 10, 0, 0)  # downward_segment_flipped
    assert s.flags & O.SEG_FLIPPED and s.y0 <= s.y1
    s = seg(5, 0, 5, 10)  # vertical_segment_dxdy_zero
    assert s.flags & O.SEG_VERT and s.dxdy == 0.0 and s.dxdy_fp == 0
    s = seg(3, 7, 3, 7)  # degenerate_point_segment
    assert s.flags & O.SEG_HORIZ and s.flags & O.SEG_VERT and not s.flags & O.SEG_FLIPPED
    s = seg(2, 1, 6, 5)  # sloped_segment_dxdy
    assert s.dxdy == 1.0 and not s.flags & (O.SEG_HORIZ | O.SEG_VERT | O.SEG_FLIPPED)
    s = seg(6, 5, 2, 1)  # flipped_sloped_segment_dxdy_consistent
    assert s.flags & O.SEG_FLIPPED and s.dxdy == 1.0 and s.y0 <= s.y1
    assert seg(0, 0, 4, 4).dxdy_fp == 65536  # dxdy_fp_matches_dxdy_for_slope_one
    assert seg(0, 0, 2, 4).dxdy_fp == 32768  # ..._half_slope
    assert seg(0, 2, 5, 2).dxdy_fp == 0 and seg(3, 0, 3, 5).dxdy_fp == 0


def test_dxdy_fp_rounds_half_away_and_saturates():
    # dxdy*65536 rounded half away from zero (f64::round), saturated to i32.
    assert seg(0, 0, 1, 2 ** 17).dxdy_fp == 1  # 0.5 -> 1
    assert seg(0, 2 ** 17, 1, 0).dxdy_fp == -1  # -0.5 -> -1 (flipped: slope from input order)
    assert seg(0, 0, 1e6, 1e-3).dxdy_fp == 2**31 - 1
    assert seg(0, 0, -1e6, 1e-3).dxdy_fp == -2**31


def test_aa_scale_multiplies_coords():
    # xpath.rs:533-550
    segs = O.add_segments(np.array([[1, 0, 3, 2]], float))
    arr = (O.Seg * 1)(segs[0])
    import ctypes
    O.lib().rro_aa_scale(ctypes.cast(arr, ctypes.c_void_p), 1)
    s = arr[0]
    assert (s.x0, s.y0, s.x1, s.y1, s.dxdy) == (4.0, 0.0, 12.0, 8.0, segs[0].dxdy)


def triangle():
    p = O.Path()
    p.move_to(0, 0)
    p.line_to(4, 0)
    p.line_to(2, 4)
    p.close(False)
    return p


def test_triangle_from_path_three_segments():
    # xpath.rs:564-576 (close() already added the closing lineTo)
    xy, fl = triangle().arrays()
    assert len(O.xpath_edges(xy, fl, flatness=1.0, close_subpaths=False)) == 3
    assert len(O.xpath_edges(xy, fl, flatness=1.0, close_subpaths=True)) == 3


# ---- crates/raster/src/path/mod.rs: builder state machine ------------------

def test_path_builder_errors_and_close():
    p = O.Path()
    assert p.line_to(1, 1) == 1  # NoCurPt
    assert p.move_to(0, 0) == 0
    assert p.move_to(1, 1) == 2  # BogusPath (one-point subpath active)
    p.line_to(5, 0)
    p.close(False)  # first != last -> lineTo(first)
    xy, fl = p.arrays()
    assert len(fl) == 3 and tuple(xy[-1]) == (0.0, 0.0)
    assert fl[0] == O.PATH_FIRST | O.PATH_CLOSED and fl[-1] == O.PATH_LAST | O.PATH_CLOSED
    assert p.line_to(1, 1) == 1  # no current point after close
    q = O.Path()
    q.move_to(0, 0)
    q.close(True)  # force adds the lineTo even for one point
    assert len(q.arrays()[1]) == 2


def test_curve_chaining_quirk():
    # xpath.rs:209-246: after a curve the LAST flag is never checked, so a
    # subpath ending with a curve exactly on its start point (no closing line,
    # path/mod.rs:431) is chained into the next subpath by a connector edge,
    # and the last subpath closes back to the first subpath's start.
    p = O.Path()
    p.move_to(0, 0)
    p.curve_to(10, 0, 10, 10, 0, 10)
    p.curve_to(-10, 10, -10, 0, 0, 0)
    p.close(False)
    p.move_to(20, 20)
    p.line_to(30, 20)
    p.line_to(30, 30)
    p.close(False)
    xy, fl = p.arrays()
    e = O.xpath_edges(xy, fl, flatness=0.1)
    starts = [tuple(r) for r in e]
    assert (0.0, 0.0, 20.0, 20.0) in starts  # connector into the second subpath
    assert tuple(e[-1]) == (20.0, 20.0, 0.0, 0.0)  # closes to the FIRST subpath start
    # with a closing line instead, the subpaths are separate
    p2 = O.Path()
    p2.move_to(0, 0)
    p2.line_to(10, 0)
    p2.line_to(0, 10)
    p2.close(False)
    p2.move_to(20, 20)
    p2.line_to(30, 20)
    p2.line_to(30, 30)
    p2.close(False)
    e2 = O.xpath_edges(*p2.arrays(), flatness=0.1)
    assert (0.0, 0.0, 20.0, 20.0) not in [tuple(r) for r in e2]


# ---- crates/raster/src/path/flatten.rs tests (:225-303) ----------------------

def test_flatten_reference_cases():
    out = O.flatten_curve([0, 0, 1, 0, 2, 0, 3, 0], 0.1)
    assert len(out) and np.all(np.abs(out[:, 1]) < 1e-10) and tuple(out[-1]) == (3.0, 0.0)
    k = 0.55228475
    out = O.flatten_curve([0, 0, 0, k, 1 - k, 1, 1, 1], 1e-4)
    assert np.allclose(out[-1], [1, 1], atol=1e-10)
    out = O.flatten_curve([0, 0, 1e8, -1e8, -1e8, 1e8, 0, 0], 0.25)
    assert len(out) <= 1024
    out = O.flatten_curve([5, 7, 5, 7, 5, 7, 5, 7], 2.2e-16)
    assert len(out) == 1 and tuple(out[0]) == (5.0, 7.0)


def test_flatten_points_are_dyadic_curve_points():
    # De Casteljau midpoint subdivision emits B(t) at dyadic t = k/1024 in
    # increasing order, and every emitted chord meets the flatness test.
    rng = np.random.default_rng(7)
    for _ in range(20):
        p = rng.uniform(-50, 50, 8)
        out = O.flatten_curve(p, 0.01)
        P = p.reshape(4, 2)
        t = np.arange(1025) / 1024.0
        B = ((1 - t) ** 3)[:, None] * P[0] + (3 * (1 - t) ** 2 * t)[:, None] * P[1] + \
            (3 * (1 - t) * t ** 2)[:, None] * P[2] + (t ** 3)[:, None] * P[3]
        idx = []
        for q in out:
            d = np.abs(B - q).max(axis=1)
            i = int(np.argmin(d))
            assert d[i] < 1e-9
            idx.append(i)
        assert idx == sorted(idx) and idx[-1] == 1024
        assert len(set(idx)) == len(idx)


def test_conic_to_cubic_is_exact_degree_elevation():
    # outline.rs:106-112: the cubic equals the quadratic at every t.
    L = O.lib()
    import ctypes
    p0, pc, p3 = np.array([1.0, 2.0]), np.array([7.0, -3.0]), np.array([4.0, 9.0])
    p1, p2 = np.zeros(2), np.zeros(2)
    L.rro_conic_to_cubic(*(ctypes.c_void_p(a.ctypes.data) for a in (p0, pc, p3, p1, p2)))
    for t in np.linspace(0, 1, 11):
        q = (1 - t) ** 2 * p0 + 2 * (1 - t) * t * pc + t ** 2 * p3
        c = (1 - t) ** 3 * p0 + 3 * (1 - t) ** 2 * t * p1 + 3 * (1 - t) * t ** 2 * p2 + t ** 3 * p3
        assert np.allclose(q, c, atol=1e-12)


# ---- crates/raster/src/scanner/mod.rs tests (:696-793) -----------------------

def test_scanner_reference_cases():
    xy, fl = triangle().arrays()
    e = O.xpath_edges(xy, fl, flatness=1.0)
    s = O.Scanner(e, False, 0, 4)
    assert s.bbox()[2] <= s.bbox()[3]
    assert s.test(2, 1)
    assert not s.test(10, 2) and not s.test(2, 5)
    y_min, y_max = s.bbox()[2], s.bbox()[3]
    assert s.row(y_min - 1) == [] and s.row(y_max + 1) == []
    assert s.row(-2**31) == [] and s.row(2**31 - 1) == []
    empty = O.Scanner(np.zeros((0, 4)), False, 0, 10)
    assert empty.bbox()[0] > empty.bbox()[1]


def test_write_intersect_merge_and_disjoint():
    # scanner/mod.rs:748-767: [0,3] then [4,7] touch (x1+1 == x0) and merge;
    # [0,2] then [5,7] stay apart.  Horizontal edges give exact entries.
    s = O.Scanner(np.array([[0.0, 1.5, 3.5, 1.5], [4.0, 1.5, 7.5, 1.5]]), False, 0, 10)
    assert s.row(1) == [(0, 7, 0)]
    s = O.Scanner(np.array([[0.0, 1.5, 2.5, 1.5], [5.0, 1.5, 7.5, 1.5]]), False, 0, 10)
    assert s.row(1) == [(0, 2, 0), (5, 7, 0)]
    # vertical edges: -1 at x=3 (downward) and +1 at x=4 (upward, FLIPPED)
    # merge and their counts sum to 0
    s = O.Scanner(np.array([[3.5, 0.5, 3.5, 1.5], [4.5, 1.5, 4.5, 0.5]]), False, 0, 10)
    assert s.row(1) == [(3, 4, 0)]
    assert s.row(0) == [(3, 4, 0)]  # first row of each edge carries count 0


def test_count_crossings():
    # scanner/mod.rs:769-793 through XPathScanner::test: a -1 crossing at x=2
    # and a +1 crossing at x=9 on row 1.
    e = np.array([[2.5, 0.5, 2.5, 1.5], [9.5, 1.5, 9.5, 0.5]])
    s = O.Scanner(e, False, 0, 3)
    assert s.row(1) == [(2, 2, -1), (9, 9, 1)]
    assert s.test(2, 1)       # on the span
    assert s.test(5, 1)       # count -1 -> inside under non-zero winding
    assert not s.test(10, 1)  # back to 0
    assert not s.test(1, 1)


# ---- crates/raster/src/bitmap.rs tests (:654-680) ----------------------------

def test_aabuf_set_span_msb_first():
    L = O.lib()
    row = np.zeros(2, np.uint8)
    L.rro_aabuf_set_span(row.ctypes.data, 16, 2, 6)
    assert list(row) == [0x3c, 0x00]
    row[:] = 0
    L.rro_aabuf_set_span(row.ctypes.data, 16, 6, 10)
    assert list(row) == [0x03, 0xc0]
    row[:] = 0
    L.rro_aabuf_set_span(row.ctypes.data, 16, 0, 16)
    assert list(row) == [0xff, 0xff]


def test_nibble_coverage_matches_popcount():
    # simd/aa_coverage.rs:51-69: high nibble = even pixel.
    rng = np.random.default_rng(3)
    buf = rng.integers(0, 256, (4, 8), dtype=np.uint8)
    shape = np.zeros(16, np.uint8)
    O.lib().rro_aa_coverage_span(buf.ctypes.data, 8, 0, 16, shape.ctypes.data)
    bits = np.unpackbits(buf, axis=1)  # MSB first
    exp = bits.reshape(4, 16, 4).sum(axis=(0, 2))
    assert list(shape) == list(exp)


# ---- compositing: crates/raster/src/pipe/aa.rs + simd/composite.rs ----------

def test_fast_path_pinned_bytes():
    # pipe/aa.rs:375-411 identity_transfer_takes_fast_path_with_pinned_bytes
    # (the Rgb8 opaque fast path, one channel at a time)
    L = O.lib()
    color = [200, 100, 50]
    shape = [(i * 17) & 255 for i in range(17)]
    dst = [(i * 13) & 255 for i in range(51)]
    for i, sh in enumerate(shape):
        a = (255 * sh + 255) >> 8
        for j in range(3):
            exp = ((255 - a) * dst[3 * i + j] + a * color[j] + 255) >> 8
            assert L.rro_composite_rgb8_fast(dst[3 * i + j], color[j], 255, sh) == exp


def test_no_alpha_plane_blend_and_full_coverage():
    # pipe/aa.rs:296-311 no_alpha_plane_uses_opaque_dst (exact no-alpha arm)
    L = O.lib()
    assert 95 <= L.rro_composite_gray8(0, 200, 255, 128) <= 105
    for d in (0, 77, 255):
        assert L.rro_composite_gray8(d, 200, 255, 255) == 200  # full coverage writes src
        assert L.rro_composite_gray8(d, 200, 255, 0) == d       # zero coverage is identity
        assert L.rro_composite_rgb8_fast(d, 200, 255, 0) == d


def test_fast_vs_exact_within_one_lsb():
    # pipe/aa.rs:559-612 fast_path_matches_general_div255_within_one_lsb
    L = O.lib()
    n_one = 0
    for color in (0, 255, 200, 1, 254, 127, 128):
        for a_in in (0, 1, 128, 255):
            for sh in range(0, 256, 3):
                for d in range(0, 256, 5):
                    f = L.rro_composite_rgb8_fast(d, color, a_in, sh)
                    e = L.rro_composite_gray8(d, color, a_in, sh)
                    assert abs(f - e) <= 1
                    n_one += f != e
    assert n_one > 0


def test_rgb_to_gray_identity_on_grey():
    # render.rs:1306-1324: BT.709 integer weights sum to 10000.
    L = O.lib()
    for v in range(256):
        assert L.rro_rgb_to_gray(v, v, v) == v
