"""Pins the image / colour half of the oracle (oracle/rr_oracle_img.c).

Known answers come from the reference's own unit tests; what they leave open is
pinned by closed forms (ramps and affine grids that the interpolation must
reproduce, a 90-degree placement whose sampling is worked out by hand) and by
invariants (convexity, masks of 0 and 255).  Each check is chosen so that a
transposed index, a swapped operand or a dropped rounding term fails it.
"""
import numpy as np
import pytest

from oracle import oracle as O


# ---- blend_u8 (crates/pdf_interp/src/renderer/page/mod.rs:2366-2395) ----

@pytest.mark.parametrize("src,dst,alpha,want", [
    (200, 50, 255, 200), (0, 255, 255, 0), (255, 0, 255, 255),   # opaque -> src
    (200, 50, 0, 50),                                            # transparent -> dst
    (255, 0, 128, 128), (0, 255, 128, 127),                      # 50 % midpoints
    (255, 255, 254, 255), (255, 255, 1, 255),                    # no overflow
])
def test_blend_u8_reference_cases(src, dst, alpha, want):
    assert O.blend_u8(src, dst, alpha) == want


def test_blend_u8_truncates_weighted_mean():
    rng = np.random.default_rng(1)
    for s, d, a in rng.integers(0, 256, (2000, 3)):
        exact = (s * a + d * (255 - a)) / 255.0
        got = O.blend_u8(int(s), int(d), int(a))
        assert got == int(np.floor(exact + 1e-9)), (s, d, a)


def test_blend_multiply_is_rounded_product():
    # blend_multiply = div255(s*d) (crates/raster/src/pipe/blend.rs:24-26);
    # div255 rounds to nearest, so the result is round(s*d/255), and 255 is the
    # identity, 0 absorbs.
    for s in range(0, 256, 3):
        for d in range(0, 256, 5):
            got = O.blend_multiply(s, d)
            assert abs(got - s * d / 255.0) <= 0.5 + 1e-9, (s, d)
    for v in range(256):
        assert O.blend_multiply(255, v) == v and O.blend_multiply(v, 255) == v
        assert O.blend_multiply(0, v) == 0


# ---- composite_rgba8 / apply_soft_mask (crates/gpu/src/composite.rs:58-109) ----

def _comp(src, dst):
    d = np.array(dst, np.uint8)
    O.composite_rgba8(np.array(src, np.uint8), d)
    return d.tolist()


def test_composite_reference_cases():
    assert _comp([200, 100, 50, 255], [10, 20, 30, 128]) == [200, 100, 50, 255]
    assert _comp([200, 100, 50, 0], [10, 20, 30, 128]) == [10, 20, 30, 128]
    r = _comp([255, 255, 255, 128], [0, 0, 0, 255])
    assert all(126 <= c <= 130 for c in r[:3]) and r[3] == 255


def test_composite_onto_opaque_is_rounded_mix():
    # With an opaque destination, a_out = 255 and each channel is
    # (s*a + d*(255-a) + 127) / 255 truncated: within half a unit of the exact mix.
    rng = np.random.default_rng(2)
    n = 5000
    src = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    dst = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    dst[:, 3] = 255
    out = dst.copy()
    O.composite_rgba8(src, out)
    a = src[:, 3:4].astype(np.float64)
    exact = (src[:, :3] * a + dst[:, :3] * (255 - a)) / 255.0
    sel = (src[:, 3] != 0)
    assert np.all(out[sel, 3] == 255)
    assert np.all(np.abs(out[sel, :3] - exact[sel]) <= 0.5 + 1e-9)
    assert np.array_equal(out[~sel], dst[~sel])


def test_composite_alpha_is_source_over():
    # a_out = a_s + a_d (1 - a_s) to within half a unit; colours stay in range.
    rng = np.random.default_rng(3)
    src = rng.integers(1, 255, (4000, 4)).astype(np.uint8)
    dst = rng.integers(0, 256, (4000, 4)).astype(np.uint8)
    out = dst.copy()
    O.composite_rgba8(src, out)
    a_s, a_d = src[:, 3] / 255.0, dst[:, 3] / 255.0
    assert np.all(np.abs(out[:, 3] / 255.0 - (a_s + a_d * (1 - a_s))) <= 0.5 / 255 + 1e-9)


@pytest.mark.parametrize("px,mask,want", [
    ([100, 150, 200, 240], [255], [240]),
    ([100, 150, 200, 200], [128], [100]),
    ([100, 150, 200, 255, 10, 20, 30, 128], [0, 0], [0, 0]),
])
def test_soft_mask_reference_cases(px, mask, want):
    p = np.array(px, np.uint8)
    O.apply_soft_mask(p, np.array(mask, np.uint8))
    assert p[3::4].tolist() == want
    orig = np.array(px, np.uint8).reshape(-1, 4)
    assert np.array_equal(p.reshape(-1, 4)[:, :3], orig[:, :3])


def test_soft_mask_rounds_product():
    a, m = np.meshgrid(np.arange(256), np.arange(256))
    px = np.zeros((a.size, 4), np.uint8)
    px[:, 3] = a.ravel()
    O.apply_soft_mask(px, m.ravel().astype(np.uint8))
    assert np.all(np.abs(px[:, 3] - a.ravel() * m.ravel() / 255.0) <= 0.5 + 1e-9)


# ---- CMYK -> RGB (crates/color/src/convert.rs:465-552; crates/gpu/src/cmyk.rs:809-846) ----

def test_cmyk_matrix_reference_cases():
    cases = [(0, 0, 0, 0, 255, 255, 255), (0, 0, 0, 255, 0, 0, 0), (255, 0, 0, 0, 0, 255, 255),
             (0, 255, 0, 0, 255, 0, 255), (0, 0, 255, 0, 255, 255, 0), (128, 0, 0, 0, 127, 255, 255),
             (0, 128, 0, 0, 255, 127, 255), (0, 0, 128, 0, 255, 255, 127),
             (128, 128, 128, 0, 127, 127, 127), (128, 128, 128, 128, 63, 63, 63),
             (200, 100, 50, 80, 38, 106, 141)]
    cmyk = np.array([c[:4] for c in cases], np.uint8)
    assert O.cmyk_matrix(cmyk).tolist() == [list(c[4:]) for c in cases]


def _corner_table(g=2):
    t = np.zeros((g, g, g, g, 3), np.uint8)  # [k][c][m][y]
    for k in range(g):
        for c in range(g):
            for m in range(g):
                for y in range(g):
                    inv_k = 255 - 255 * k
                    t[k, c, m, y] = [((255 - 255 * c) * inv_k) // 255, ((255 - 255 * m) * inv_k) // 255,
                                     ((255 - 255 * y) * inv_k) // 255]
    return t.ravel()


def test_icc_clut_identity_corners():
    t = _corner_table()
    assert O.icc_clut([0, 0, 0, 0], t, 2).tolist() == [[255, 255, 255]]
    assert O.icc_clut([0, 0, 0, 255], t, 2).tolist() == [[0, 0, 0]]
    # every corner of the cube reproduces its node
    for bits in range(16):
        c, m, y, k = ((bits >> s) & 1 for s in range(4))
        got = O.icc_clut([255 * c, 255 * m, 255 * y, 255 * k], t, 2)[0]
        want = t.reshape(2, 2, 2, 2, 3)[k, c, m, y]
        assert got.tolist() == want.tolist(), bits


@pytest.mark.parametrize("axes", [(0, 1, 2), (3, 0, 1), (2, 3, 0), (1, 2, 3)])
def test_icc_clut_reproduces_axis_ramps(axes):
    # Node value 15*i along one input axis (g = 18, so 17 steps of 15 span 0..255):
    # quadrilinear interpolation of a linear function is that function, so the
    # output equals the input byte of the chosen axis.  Each output channel
    # follows a different axis; a transposed grid index fails.
    g = 18
    t = np.zeros((g, g, g, g, 3), np.uint8)   # [k][c][m][y]
    idx = np.indices((g, g, g, g))            # k, c, m, y
    axis_of = {0: idx[1], 1: idx[2], 2: idx[3], 3: idx[0]}   # input channel c, m, y, k
    for ch, ax in enumerate(axes):
        t[..., ch] = 15 * axis_of[ax]
    rng = np.random.default_rng(4)
    cmyk = rng.integers(0, 256, (3000, 4)).astype(np.uint8)
    out = O.icc_clut(cmyk, t.ravel(), g).astype(np.int32)
    want = cmyk[:, list(axes)].astype(np.int32)
    assert np.abs(out - want).max() <= 0, np.argwhere(out != want)[:5]


def test_icc_clut_g2_is_multilinear():
    # With g = 2 the result is the multilinear blend of the 16 corners; compare
    # with a float64 evaluation (allowing one unit for f32 rounding near .5).
    rng = np.random.default_rng(5)
    t = rng.integers(0, 256, 16 * 3).astype(np.uint8)
    T = t.reshape(2, 2, 2, 2, 3).astype(np.float64)
    cmyk = rng.integers(0, 256, (2000, 4)).astype(np.uint8)
    out = O.icc_clut(cmyk, t, 2).astype(np.float64)
    w = cmyk / 255.0
    ref = np.zeros((len(cmyk), 3))
    for bits in range(16):
        c, m, y, k = ((bits >> s) & 1 for s in range(4))
        wt = ((w[:, 0] if c else 1 - w[:, 0]) * (w[:, 1] if m else 1 - w[:, 1]) *
              (w[:, 2] if y else 1 - w[:, 2]) * (w[:, 3] if k else 1 - w[:, 3]))
        ref += wt[:, None] * T[k, c, m, y]
    assert np.abs(out - ref).max() <= 0.5 + 1e-3


# ---- grey CLUT (own definition: integer simplex) ----

def _grid(g, dims, fn):
    idx = np.indices((g,) * dims)
    if dims == 4:   # storage order k, c, m, y ; fn takes (c, m, y, k)
        vals = fn(idx[1], idx[2], idx[3], idx[0])
    else:           # storage order r, g, b
        vals = fn(idx[0], idx[1], idx[2])
    return np.ascontiguousarray(vals, np.uint16).ravel()


@pytest.mark.parametrize("dims", [3, 4])
@pytest.mark.parametrize("axis", [0, 1, 2, 3])
def test_clut_gray_reproduces_axis_ramp(dims, axis):
    # node = 257 * 15 * i on one axis (g = 18): the interpolant of a linear
    # function is exact, and (257*255*v + 32767) / 65535 = v.
    if axis >= dims:
        pytest.skip("axis not present")
    g = 18
    grid = _grid(g, dims, lambda *ax: 257 * 15 * ax[axis])
    rng = np.random.default_rng(10 + axis)
    for px in rng.integers(0, 256, (400, dims)).astype(np.uint8):
        assert O.clut_gray(px, grid, g) == px[axis], px


@pytest.mark.parametrize("dims", [3, 4])
def test_clut_gray_reproduces_affine_mean(dims):
    # node = 257 * 5 * (sum of indices) with g = 18 over 3 axes -> output is the
    # mean of those three channels, rounded half up: round((a+b+c)/3).
    g = 18
    grid = _grid(g, dims, lambda *ax: 257 * 5 * (ax[0] + ax[1] + ax[2]))
    rng = np.random.default_rng(20 + dims)
    for px in rng.integers(0, 256, (500, dims)).astype(np.uint8):
        s = int(px[0]) + int(px[1]) + int(px[2])
        assert O.clut_gray(px, grid, g) == (s + 1) // 3, px


@pytest.mark.parametrize("dims,g", [(3, 33), (4, 17), (4, 2), (3, 52)])
def test_clut_gray_exact_at_nodes_and_convex(dims, g):
    # Output scale: weights sum to 255 and (acc + 32767) / 65535 divides by
    # 255 * 257, so a node value n reads back as round-half-up(n / 257).
    rng = np.random.default_rng(30 + g)
    grid = rng.integers(0, 65536, g ** dims).astype(np.uint16)
    G = grid.reshape((g,) * dims)

    def node(ii):  # ii in channel order (c, m, y, k) / (r, g, b)
        return int(G[(ii[3], ii[0], ii[1], ii[2]) if dims == 4 else tuple(ii)])

    def scale(n):
        return (n * 255 + 32767) // 65535

    for px in rng.integers(0, 256, (600, dims)).astype(np.uint8):
        t = px.astype(np.int64) * (g - 1)
        i0 = t // 255
        i1 = np.minimum(i0 + 1, g - 1)
        corners = [node([i1[a] if (bits >> a) & 1 else i0[a] for a in range(dims)])
                   for bits in range(1 << dims)]
        got = O.clut_gray(px, grid, g)
        assert scale(min(corners)) <= got <= scale(max(corners)), px
    node_vals = [v for v in range(256) if (v * (g - 1)) % 255 == 0]
    for _ in range(100):
        px = np.array(rng.choice(node_vals, dims), np.uint8)
        assert O.clut_gray(px, grid, g) == scale(node(px.astype(np.int64) * (g - 1) // 255))


def test_clut_gray_is_simplex_not_multilinear():
    # On one cell with corners 0 except the far corner (all axes +1) the Kuhn
    # simplex interpolant is min(f) over the axes, not the product of the
    # fractions: at fractions (1/2, 1/2, 1/2) simplex gives 1/2, trilinear 1/8.
    g = 2
    grid = np.zeros(8, np.uint16)
    grid[7] = 65535
    assert O.clut_gray(np.array([128, 128, 128], np.uint8), grid, g) == 128
    assert O.clut_gray(np.array([255, 64, 200], np.uint8), grid, g) == 64


# ---- image placement (crates/pdf_interp/src/renderer/page/mod.rs:1600-1676,1899-1974) ----

def _ramp_grid3():
    return _grid(18, 3, lambda r, g, b: 257 * 15 * r)   # grey = R channel


@pytest.mark.parametrize("k,n,flip", [(2, 4, False), (1, 2, False), (4, 1, True), (2, 2, True)])
def test_place_image_axis_aligned_closed_form(k, n, flip):
    # blit_image's axis-aligned path (page/mod.rs:1767-1898) for a CTM
    # (+-k W, 0, 0, n H, e, f) with W, H, k, n powers of two and integer e, f: the
    # Q32 column stepping and the f64 row mapping are exact, so the path is a
    # nearest-neighbour scale: iy = (dy - by0) // n and, unflipped,
    # ix = (dx - e) // k; x-flipped (a < 0, u = 1 at the bbox's left edge),
    # ix = min(W - 1, W - ceil((dx - bx0) / k)).
    W, H, ph, pw = 16, 16, 120, 100
    e, f = (70.0 if flip else 9.0), 20.0
    rng = np.random.default_rng(39 + k)
    px = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    canvas = np.full((ph, pw), 7, np.uint8)
    a = -k * W if flip else k * W
    got_n = O.place_image(canvas, px, None, (float(a), 0.0, 0.0, float(n * H), e, f), 0, _ramp_grid3(), 18)
    want = np.full((ph, pw), 7, np.uint8)
    by0, by1 = int(ph - f - n * H), int(ph - f)
    bx0 = int(e - k * W) if flip else int(e)
    for dy in range(by0, by1):
        for dx in range(bx0, bx0 + k * W):
            iy = (dy - by0) // n
            j = dx - bx0
            ix = min(W - 1, W - (j + k - 1) // k) if flip else j // k
            want[dy, dx] = px[iy, ix, 0]
    assert got_n == k * W * n * H
    assert np.array_equal(canvas, want)


def test_place_image_quarter_turn_closed_form():
    # CTM (0, W, -H, 0, e, f): image x runs up the page, image y runs right.
    # With W, H powers of two every f64 step is exact, so the general path's
    # corner sampling gives ix = min(page_h - f - dy, W - 1), iy = H + dx - e.
    W, H, ph, pw = 64, 32, 200, 160
    e, f = 100.0, 50.0
    rng = np.random.default_rng(40)
    px = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    canvas = np.full((ph, pw), 7, np.uint8)
    n = O.place_image(canvas, px, None, (0.0, float(W), -float(H), 0.0, e, f), 0, _ramp_grid3(), 18)
    want = np.full((ph, pw), 7, np.uint8)
    for dy in range(int(ph - f - W), int(ph - f)):
        for dx in range(int(e - H), int(e)):
            ix = min(int(ph - f - dy), W - 1)
            iy = H + dx - int(e)
            want[dy, dx] = px[iy, ix, 0]
    assert n == W * H
    assert np.array_equal(canvas, want)


def test_place_image_mask_and_blend_modes():
    W, H, ph, pw = 64, 32, 200, 160
    ctm = (0.0, float(W), -float(H), 0.0, 100.0, 50.0)
    rng = np.random.default_rng(41)
    px = rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
    grid = _ramp_grid3()
    base = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
    # mask 0 everywhere: nothing written
    c0 = base.copy()
    assert O.place_image(c0, px, np.zeros((H, W), np.uint8), ctm, 0, grid, 18) == 0
    assert np.array_equal(c0, base)
    # opaque normal vs multiply vs masked: per pixel, from the normal result
    cn = base.copy()
    O.place_image(cn, px, None, ctm, 0, grid, 18)
    cm = base.copy()
    O.place_image(cm, px, None, ctm, 1, grid, 18)
    mask = rng.integers(1, 256, (H, W)).astype(np.uint8)
    ch = base.copy()
    O.place_image(ch, px, mask, ctm, 0, grid, 18)
    touched = cn != base
    ys, xs = np.nonzero(np.ones_like(base, bool))
    inside = np.zeros_like(base, bool)
    inside[int(ph - 50 - W):int(ph - 50), 100 - H:100] = True
    for y, x in zip(*np.nonzero(inside)):
        s, d = int(cn[y, x]), int(base[y, x])
        assert cm[y, x] == O.blend_multiply(s, d)
        ix = min(int(ph - 50 - y), W - 1)
        iy = H + x - 100
        a = int(mask[iy, ix])
        assert ch[y, x] == (s if a == 255 else O.blend_u8(s, d, a))
    assert np.array_equal(cn[~inside], base[~inside])
    assert not np.any(touched & ~inside)


def test_place_image_bbox_clamps_to_canvas():
    # An image hanging over the canvas edges writes exactly the clamped bbox.
    W, H = 64, 32
    canvas = np.full((40, 40), 9, np.uint8)
    px = np.full((H, W, 3), 255, np.uint8)
    grid = _ramp_grid3()
    n = O.place_image(canvas, px, None, (0.0, float(W), -float(H), 0.0, 20.0, 10.0), 0, grid, 18)
    # x in [20-32, 20) -> [0, 20); y_pdf in [10, 74] -> dy in [40-74, 30) -> [0, 30)
    assert n == 20 * 30
    assert np.all(canvas[:30, :20] == 255) and np.all(canvas[30:, :] == 9) and np.all(canvas[:, 20:] == 9)
