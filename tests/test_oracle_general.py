"""Pins the oracle's general pipe (soft mask, Multiply; no alpha plane,
crates/raster/src/pipe/general.rs:270-344, blend.rs:24-26) and its Rgb8
canvas, against the reference's own known answers and closed forms."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib


def div255(x):
    # crates/color/src/convert.rs:50-57 (pinned by its own reference cases)
    x = np.asarray(x, np.int64)
    return np.minimum((x + (x >> 8) + 0x80) >> 8, 255)


# ---- the reference's known answers (pipe/general.rs:518-545) ----

def test_compute_a_src_reference_cases():
    f = L().rro_compute_a_src
    assert f(200, 0xFF, 0xFF, 0, 0) == 200
    assert f(255, 0xFF, 0, 0, 1) == 0
    assert f(255, 128, 0xFF, 1, 0) == L().rro_div255(255 * 128)
    assert f(255, 128, 128, 1, 1) == L().rro_div255(L().rro_div255(255 * 128) * 128)


def test_general_no_alpha_readings_of_reference_cases():
    # general.rs:572-611 and 637-651 run the alpha-plane arm; the fill's
    # bitmaps have no alpha plane (implicit a_dst = 255), where the same
    # inputs give: opaque source -> source; a_input 0 -> destination;
    # Multiply(128, 200) at full coverage = div255(128 * 200) ~ 100.
    g = L().rro_composite_general
    for d in (10, 20, 30, 200):
        assert g(d, 200, 255, 255, 255, 0, 0) == 200
        assert g(d, 255, 0, 255, 255, 0, 0) == d
    v = g(200, 128, 255, 255, 255, 0, 1)
    assert abs(v - 100) <= 1 and v == div255(128 * 200)
    # soft mask 128, white over black, full coverage: a_src = div255(255 * 128)
    assert g(0, 255, 255, 255, 128, 1, 0) == L().rro_div255(255 * 128) == 128


# ---- closed forms ----

@pytest.fixture(scope="module")
def samples():
    rng = np.random.default_rng(7)
    s = rng.integers(0, 256, (20000, 5))
    edge = np.array([0, 1, 127, 128, 254, 255])
    s[:3000] = edge[rng.integers(0, len(edge), (3000, 5))]
    return [tuple(int(v) for v in r) for r in s]


def test_normal_without_mask_equals_aa_gray8(samples):
    # The general pipe with Normal and no soft mask is the AA path's Gray8
    # formula: a = div255(a_in * shape), c = div255((255 - a) d + a s).
    g, aa = L().rro_composite_general, L().rro_composite_gray8
    for d, src, a_in, sh, sm in samples:
        assert g(d, src, a_in, sh, sm, 0, 0) == aa(d, src, a_in, sh)


def test_mask_extremes_and_zero_shape(samples):
    g = L().rro_composite_general
    for d, src, a_in, sh, sm in samples:
        for mult in (0, 1):
            assert g(d, src, a_in, sh, 255, 1, mult) == g(d, src, a_in, sh, sm, 0, mult)
            assert g(d, src, a_in, sh, 0, 1, mult) == d
            assert g(d, src, a_in, 0, sm, 1, mult) == d
            assert g(d, src, a_in, 0, sm, 0, mult) == d


def test_multiply_closed_forms(samples):
    g = L().rro_composite_general
    for d, src, a_in, sh, sm in samples:
        assert g(255, src, a_in, sh, sm, 1, 1) == g(255, src, a_in, sh, sm, 1, 0)  # white: Multiply = Normal
        assert g(0, src, a_in, sh, sm, 1, 1) == 0                                  # black stays black
        assert g(d, src, 255, 255, 255, 0, 1) == div255(src * d)                   # opaque full cover
        lo, hi = sorted((d, int(div255(src * d))))
        assert lo <= g(d, src, a_in, sh, sm, 1, 1) <= hi                           # a convex mix


# ---- fills ----

def _rect(x0, y0, x1, y1):
    return np.array([[x0, y0, x1, y0], [x1, y0, x1, y1], [x1, y1, x0, y1], [x0, y1, x0, y0]], np.float64)


def _tri_star(rng, W, H):
    pts = np.stack([rng.uniform(-2, W + 2, 7), rng.uniform(-2, H + 2, 7)], 1)
    return np.array([[*pts[i], *pts[(i + 3) % 7]] for i in range(7)], np.float64)


def test_rgb8_canvas_is_three_grey_planes():
    # A colour fill on the Rgb8 canvas equals, channel by channel, the same
    # fill of that channel's value on a grey canvas of the plane: the AA path
    # with the Rgb8 fast composite, the general pipe with its div255 form.
    rng = np.random.default_rng(11)
    W, H = 41, 33
    base = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    soft = rng.integers(0, 256, (H, W)).astype(np.uint8)
    for k in range(24):
        e = _tri_star(rng, W, H)
        flags = int(rng.integers(0, 8))
        col = tuple(int(v) for v in rng.integers(0, 256, 3))
        a_in = int(rng.choice([255, int(rng.integers(0, 256))]))
        mode = k % 2
        rgb = base.copy()
        O.fill_ex(rgb, e, flags, col, a_in, mode, O.FLAVOUR_RGB8, soft_mask=soft)
        for ch in range(3):
            plane = np.ascontiguousarray(base[:, :, ch])
            O.fill_ex(plane, e, flags, col[ch], a_in, mode, O.FLAVOUR_RGB8_FAST, soft_mask=soft)
            assert np.array_equal(rgb[:, :, ch], plane), f"case {k} channel {ch}"


def test_soft_mask_extremes_on_fills():
    rng = np.random.default_rng(12)
    W, H = 37, 29
    base = rng.integers(0, 256, (H, W)).astype(np.uint8)
    for k in range(12):
        e = _tri_star(rng, W, H)
        eo = k % 2
        a_in = int(rng.integers(1, 256))
        plain = base.copy()
        O.fill_ex(plain, e, eo, 90, a_in)
        full = base.copy()
        O.fill_ex(full, e, eo | O.FILL_SOFT_MASK, 90, a_in, soft_mask=np.full((H, W), 255, np.uint8))
        assert np.array_equal(full, plain)      # mask 255 everywhere: the AA path's Gray8 pixels
        none = base.copy()
        O.fill_ex(none, e, eo | O.FILL_SOFT_MASK | (k & 2), 90, a_in, soft_mask=np.zeros((H, W), np.uint8))
        assert np.array_equal(none, base)       # mask 0: nothing painted


def test_multiply_rectangle_closed_form():
    # An opaque rectangle with Multiply over a random page: interior pixels
    # (coverage 16) become blend_multiply = div255(s * d), pixels outside stay.
    rng = np.random.default_rng(13)
    W, H = 30, 24
    base = rng.integers(0, 256, (H, W)).astype(np.uint8)
    out = base.copy()
    _, _, shape = O.fill_ex(out, _rect(5.0, 4.0, 21.0, 17.0), O.FILL_MULTIPLY, 77, want_shape=True)
    inner = shape == 255
    assert inner[6:16, 6:20].all()
    assert np.array_equal(out[inner], div255(77 * base[inner].astype(np.int64)).astype(np.uint8))
    assert np.array_equal(out[shape == 0], base[shape == 0])


def test_rgb_to_gray_page_identity_and_weights():
    rng = np.random.default_rng(14)
    v = rng.integers(0, 256, (9, 13)).astype(np.uint8)
    assert np.array_equal(O.rgb_to_gray_page(np.repeat(v[:, :, None], 3, 2)), v)
    rgb = rng.integers(0, 256, (9, 13, 3)).astype(np.int64)
    want = (2126 * rgb[..., 0] + 7152 * rgb[..., 1] + 722 * rgb[..., 2] + 5000) // 10000
    assert np.array_equal(O.rgb_to_gray_page(rgb.astype(np.uint8)), want.astype(np.uint8))


# ---- planted mistakes (RRO_MUTANT in oracle/rr_oracle.c) must be rejected ----

@pytest.fixture(scope="module")
def mutant_builds():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["make", "-s", "-C", os.path.join(root, "oracle"), "mutants"], check=True)


@pytest.mark.parametrize("m,check", [(7, test_mask_extremes_and_zero_shape), (8, test_multiply_closed_forms)])
def test_planted_mistake_is_rejected(mutant_builds, samples, m, check):
    with O.using(O.mutant_path(m)):
        with pytest.raises(AssertionError):
            check(samples)


# ---- the static dispatch rule's bbox (gpu_ops.rs:52-84) ----

@pytest.mark.parametrize("edges,want", [
    ([[0, 0, 16, 0], [16, 0, 16, 16], [16, 16, 0, 0]], 256),          # 16 x 16 exactly: the threshold
    ([[0, 0, 15, 0], [15, 0, 15, 17], [15, 17, 0, 0]], 255),          # 15 x 17: below it
    ([[0.5, 0.5, 15.2, 0.5], [15.2, 0.5, 0.5, 15.9]], 256),           # floor / ceil to whole pixels
    ([[-40, -3, 10, -3], [10, -3, 10, 12], [10, 12, -40, -3]], 120),  # clamped at 0: 10 x 12
    ([[95, 70, 130, 70], [130, 70, 95, 99]], 50),                     # clamped at the 100 x 80 page: 5 x 10
    ([[3, 3, float("nan"), 9], [3, 3, 20, 19]], 17 * 16),            # NaN skipped by f64::min / max
    ([[3, 3, float("inf"), 9], [3, 3, 20, 19]], -1),                 # not finite: CPU
    ([[200, 3, 230, 9]], -1),                                         # right of the page: empty
])
def test_ref_bbox_area_cases(edges, want):
    assert O.ref_bbox_area(np.array(edges, np.float64), 100, 80) == want
