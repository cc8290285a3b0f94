"""Image placement + soft mask + Normal/Multiply through the grey CLUT, and the
reference's elementwise GPU ops (ICC CLUT / matrix, RGBA composite, soft mask),
on the device through the C-ABI, bit-exact against the oracle."""
import ctypes

import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _oracle_pages(iset, n_pages=None, background=255):
    n_pages = iset.n_pages if n_pages is None else n_pages
    canv = [np.full((iset.page_h, iset.page_w), background, np.uint8) for _ in range(n_pages)]
    for im in iset.images:
        grid, g = (iset.grid4, iset.grid4_n) if im.channels == 4 else (iset.grid3, iset.grid3_n)
        assert O.place_image(canv[im.page], im.pixels, im.smask, im.ctm, im.blend, grid, g) >= 0
    return canv


def _compare(got, want, what):
    if not np.array_equal(got, want):
        ys, xs = np.nonzero(got != want)
        raise AssertionError(f"{what}: {len(ys)} px differ, first y={ys[0]} x={xs[0]} "
                             f"got {got[ys[0], xs[0]]} want {want[ys[0], xs[0]]}")


def test_image_workload_pages(backend):
    iset = R.images(7, 8, 4)   # CMYK/Normal and RGB/Multiply alternate, 2 pages
    di = R.DeviceImages(backend, iset)
    try:
        di.run()
        got = [di.read_canvas(i) for i in range(iset.n_pages)]
    finally:
        di.free()
    for i, (g, w) in enumerate(zip(got, _oracle_pages(iset))):
        _compare(g, w, f"page {i}")


def _single(be, px, smask, ctm, blend, grid, g, pw, ph, background, grid_shift=0):
    """Place one hand-made image on a pw x ph canvas on the device; grid_shift
    nodes of padding in front of the grid move its device address off 16-byte
    alignment."""
    grid = np.concatenate([np.zeros(grid_shift, np.uint16), grid])
    iset = R.ImageSet([R.Image(px, smask, px.shape[2], blend, ctm, 0)], 1, pw, ph,
                      grid if px.shape[2] == 4 else np.zeros(16, np.uint16), g if px.shape[2] == 4 else 2,
                      grid if px.shape[2] == 3 else np.zeros(8, np.uint16), g if px.shape[2] == 3 else 2)
    di = R.DeviceImages(be, iset)
    try:
        di.reset(0)
        be.upload(di.canvases[0], np.pad(background, ((0, 0), (0, di.stride - pw))))
        b = di.batch()
        if px.shape[2] == 4:
            b.grid4 += 2 * grid_shift
        else:
            b.grid3 += 2 * grid_shift
        be.begin_page()
        be.record_place_images(b)
        be.wait_page(be.submit_page())
        return di.read_canvas(0)
    finally:
        di.free()


@pytest.mark.parametrize("channels,blend,masked", [(3, 0, False), (3, 1, True), (4, 0, True), (4, 1, False)])
def test_quarter_turn_and_ragged_edges(backend, channels, blend, masked):
    # Canvas width 157 (not a multiple of 4) and a bbox starting at x = 61, so
    # the dword read-modify-write covers partial words on both sides.
    rng = np.random.default_rng(50 + channels + 2 * blend)
    W, H, pw, ph = 64, 32, 157, 120
    px = rng.integers(0, 256, (H, W, channels)).astype(np.uint8)
    sm = rng.integers(0, 256, (H, W)).astype(np.uint8) if masked else None
    g = 17 if channels == 4 else 33
    grid = rng.integers(0, 65536, g ** channels).astype(np.uint16)
    ctm = (0.0, float(W), -float(H), 0.0, 93.0, 21.0)
    bg = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
    got = _single(backend, px, sm, ctm, blend, grid, g, pw, ph, bg)
    want = bg.copy()
    O.place_image(want, px, sm, ctm, blend, grid, g)
    _compare(got, want, "quarter turn")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_small_affine(backend, seed):
    # Small images under random rotations/scales, partly off the canvas.
    rng = np.random.default_rng(100 + seed)
    pw, ph = 301, 233
    for k in range(6):
        ch = 4 if k % 2 == 0 else 3
        W, H = (int(v) for v in rng.integers(5, 90, 2))
        px = rng.integers(0, 256, (H, W, ch)).astype(np.uint8)
        sm = rng.integers(0, 256, (H, W)).astype(np.uint8) if k % 3 else None
        th = np.deg2rad(rng.uniform(3, 80))
        s = rng.uniform(0.5, 3.0)
        a, b, c, d = s * W * np.cos(th), s * W * np.sin(th), -s * H * np.sin(th), s * H * np.cos(th)
        e, f = rng.uniform(-50, pw), rng.uniform(-50, ph)
        ctm = (a, b, c, d, e, f)
        g = 17 if ch == 4 else 33
        grid = rng.integers(0, 65536, g ** ch).astype(np.uint16)
        bg = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
        got = _single(backend, px, sm, ctm, k % 2, grid, g, pw, ph, bg)
        want = bg.copy()
        assert O.place_image(want, px, sm, ctm, k % 2, grid, g) >= 0
        _compare(got, want, f"seed {seed} image {k}")


@pytest.mark.parametrize("scale", [1.7, 0.6])
@pytest.mark.parametrize("channels,g,shift", [(4, 40, 0), (3, 64, 0), (4, 17, 1), (3, 33, 3)])
def test_grid_beyond_lds_and_unaligned_grid(backend, channels, g, shift, scale):
    # 40^4 and 64^3 u16 grids are larger than a block's LDS, so taps past the
    # LDS copy read global memory; shifted grids take the 2-byte copy loop.
    # Scale 1.7 enlarges the image (grey + alpha converted ahead per image
    # pixel), scale 0.6 shrinks it (grid interpolation per canvas pixel).
    rng = np.random.default_rng(70 + g + shift)
    W, H, pw, ph = 120, 90, 260, 230
    px = rng.integers(0, 256, (H, W, channels)).astype(np.uint8)
    sm = rng.integers(0, 256, (H, W)).astype(np.uint8)
    grid = rng.integers(0, 65536, g ** channels).astype(np.uint16)
    th, s = np.deg2rad(20.0), scale
    ctm = (s * W * np.cos(th), s * W * np.sin(th), -s * H * np.sin(th), s * H * np.cos(th), 70.0, 15.0)
    bg = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
    got = _single(backend, px, sm, ctm, int(channels == 3), grid, g, pw, ph, bg, grid_shift=shift)
    want = bg.copy()
    assert O.place_image(want, px, sm, ctm, int(channels == 3), grid, g) >= 0
    _compare(got, want, f"g {g} shift {shift}")


@pytest.mark.parametrize("channels", [3, 4])
def test_grid_corner_colours(backend, channels):
    # Channel bytes at the grid's first and last nodes (0, 255) and next to
    # cell boundaries: vertices of weight zero, steps off the grid's last node.
    rng = np.random.default_rng(80 + channels)
    W, H, pw, ph = 48, 40, 150, 140
    vals = np.array([0, 1, 15, 16, 128, 239, 240, 254, 255], np.uint8)
    px = vals[rng.integers(0, len(vals), (H, W, channels))]
    g = 17 if channels == 4 else 33
    grid = rng.integers(0, 65536, g ** channels).astype(np.uint16)
    th, s = np.deg2rad(12.0), 2.5
    ctm = (s * W * np.cos(th), s * W * np.sin(th), -s * H * np.sin(th), s * H * np.cos(th), 20.0, 5.0)
    bg = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
    got = _single(backend, px, None, ctm, 0, grid, g, pw, ph, bg)
    want = bg.copy()
    assert O.place_image(want, px, None, ctm, 0, grid, g) >= 0
    _compare(got, want, "corner colours")


@pytest.mark.parametrize("case", range(6))
def test_axis_aligned_ctm(backend, case):
    # CTMs without rotation take blit_image's axis-aligned path (Q32 column
    # stepping, page/mod.rs:1767-1898): enlarged, shrunk, x-flipped (a < 0),
    # y-flipped (d < 0), partly off the canvas, and a sub-threshold shear.
    rng = np.random.default_rng(140 + case)
    pw, ph = 203, 171
    ch = 4 if case % 2 == 0 else 3
    W, H = (int(v) for v in rng.integers(6, 70, 2))
    a, d = [(3.1, 2.2), (0.45, 0.7), (-2.5, 1.3), (1.7, -2.9), (5.0, 4.0), (1.3, 1.1)][case]
    a, d = a * W, d * H
    b = c = 0.0
    if case == 5:
        b, c = 0.4 / W, -0.3 / H  # |b| w < 0.5 and |c| h < 0.5: still axis-aligned
    e, f = [(20.0, 15.0), (31.5, 40.25), (180.0, 9.0), (12.0, 160.0), (-60.0, -30.0), (44.0, 13.0)][case]
    ctm = (a, b, c, d, e, f)
    px = rng.integers(0, 256, (H, W, ch)).astype(np.uint8)
    sm = rng.integers(0, 256, (H, W)).astype(np.uint8) if case % 3 else None
    g = 17 if ch == 4 else 33
    grid = rng.integers(0, 65536, g ** ch).astype(np.uint16)
    bg = rng.integers(0, 256, (ph, pw)).astype(np.uint8)
    want = bg.copy()
    assert O.place_image(want, px, sm, ctm, case % 2, grid, g) > 0
    got = _single(backend, px, sm, ctm, case % 2, grid, g, pw, ph, bg)
    _compare(got, want, f"axis-aligned case {case}")


def test_rejected_batch_leaves_canvas_unchanged(backend):
    # Image 1 names a page that does not exist: the whole batch is refused
    # before image 0 is drawn, so a CPU fallback applies every image once.
    rng = np.random.default_rng(150)
    pw, ph = 96, 80
    px = rng.integers(0, 256, (30, 40, 3)).astype(np.uint8)
    ctm = (70.0, 8.0, -6.0, 50.0, 10.0, 10.0)
    iset = R.ImageSet([R.Image(px, None, 3, 1, ctm, 0), R.Image(px, None, 3, 0, ctm, 0)], 1, pw, ph,
                      np.zeros(16, np.uint16), 2, rng.integers(0, 65536, 33 ** 3).astype(np.uint16), 33)
    di = R.DeviceImages(backend, iset)
    try:
        di.reset(77)
        b = di.batch()
        descs = ctypes.cast(b.images, ctypes.POINTER(R.ImageDesc))
        descs[1].page = 5
        backend.begin_page()
        with pytest.raises(R.RrError) as ei:
            backend.record_place_images(b)
        assert "page index" in str(ei.value)
        backend.wait_page(backend.submit_page())
        assert np.all(di.read_canvas(0) == 77)
    finally:
        di.free()


# ---- elementwise reference ops ----

def _run(be, fn):
    be.begin_page()
    fn()
    be.wait_page(be.submit_page())


@pytest.mark.parametrize("g", [2, 17, 33])
def test_icc_clut(backend, g):
    rng = np.random.default_rng(60 + g)
    n = 100_003
    cmyk = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    table = rng.integers(0, 256, g ** 4 * 3).astype(np.uint8)
    d_in, d_t = backend.to_device(cmyk), backend.to_device(table)
    d_out = backend.alloc(n * 3)
    try:
        _run(backend, lambda: backend.record_icc_clut(d_in, d_out, d_t, table.nbytes, n))
        got = backend.download(d_out, np.zeros((n, 3), np.uint8))
    finally:
        for b in (d_in, d_t, d_out):
            backend.free(b)
    want = O.icc_clut(cmyk, table, g)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


@pytest.mark.parametrize("nbytes", [0, 3, 17 ** 4 * 3 + 3, 17 ** 4 * 3 - 3, 256 ** 4 * 3])
def test_icc_clut_rejects_bad_table_length(backend, nbytes):
    # grid_n_from_clut_len (params.rs:127-147): only g^4 * 3 bytes with
    # 2 <= g <= 255 name a table; anything else is refused before a launch.
    d_t = backend.alloc(64)
    try:
        backend.begin_page()
        with pytest.raises(R.RrError) as ei:
            backend.record_icc_clut(d_t, d_t, d_t, nbytes, 16)
        assert ei.value.code == R.RR_E_INVARIANT
        backend.wait_page(backend.submit_page())
    finally:
        backend.free(d_t)


def test_cmyk_matrix(backend):
    a = np.arange(256, dtype=np.uint8)
    rng = np.random.default_rng(61)
    cmyk = np.concatenate([np.stack([a, a[::-1], a, a[::-1]], 1),
                           rng.integers(0, 256, (77_777, 4)).astype(np.uint8)])
    n = len(cmyk)
    d_in, d_out = backend.to_device(cmyk), backend.alloc(n * 3)
    try:
        _run(backend, lambda: backend.record_icc_clut(d_in, d_out, None, 0, n))
        got = backend.download(d_out, np.zeros((n, 3), np.uint8))
    finally:
        backend.free(d_in)
        backend.free(d_out)
    assert np.array_equal(got, O.cmyk_matrix(cmyk))


def test_composite_and_soft_mask(backend):
    rng = np.random.default_rng(62)
    n = 123_457
    src = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    src[::7, 3] = 0
    src[1::7, 3] = 255
    dst = rng.integers(0, 256, (n, 4)).astype(np.uint8)
    mask = rng.integers(0, 256, n).astype(np.uint8)
    d_src, d_dst, d_m = backend.to_device(src), backend.to_device(dst), backend.to_device(mask)
    try:
        _run(backend, lambda: backend.record_composite(d_src, d_dst, n))
        got = backend.download(d_dst, np.zeros((n, 4), np.uint8))
        _run(backend, lambda: backend.record_apply_soft_mask(d_dst, d_m, n))
        got2 = backend.download(d_dst, np.zeros((n, 4), np.uint8))
    finally:
        for b in (d_src, d_dst, d_m):
            backend.free(b)
    want = dst.copy()
    O.composite_rgba8(src, want)
    assert np.array_equal(got, want)
    O.apply_soft_mask(want, mask)
    assert np.array_equal(got2, want)


def test_elementwise_empty_is_noop(backend):
    _run(backend, lambda: (backend.record_composite(None, None, 0),
                           backend.record_apply_soft_mask(None, None, 0),
                           backend.record_icc_clut(None, None, None, 0, 0)))


@pytest.mark.parametrize("width,rgb_pad,gray_pad", [(1024, 0, 0), (1001, 1, 3), (7, 5, 1)])
def test_rgb_to_gray(backend, width, rgb_pad, gray_pad):
    # Aligned rows (dword path) and ragged ones (byte path); padding bytes of
    # the grey bitmap stay untouched; grey pixels map to themselves.
    rng = np.random.default_rng(90 + width)
    H = 37
    rs, gs = 3 * width + rgb_pad, width + gray_pad
    rgb = rng.integers(0, 256, (H, rs)).astype(np.uint8)
    v = rng.integers(0, 256, width).astype(np.uint8)
    rgb[5, :3 * width] = np.repeat(v, 3)  # a grey row
    d_rgb = backend.to_device(rgb)
    d_gray = backend.to_device(np.full((H, gs), 77, np.uint8))
    try:
        _run(backend, lambda: backend.record_rgb_to_gray(d_rgb, rs, d_gray, gs, width, H))
        got = backend.download(d_gray, np.zeros((H, gs), np.uint8))
    finally:
        backend.free(d_rgb)
        backend.free(d_gray)
    L = O.lib()
    want = np.full((H, gs), 77, np.uint8)
    for y in range(H):
        for x in range(width):
            r, g, b = (int(c) for c in rgb[y, 3 * x:3 * x + 3])
            want[y, x] = L.rro_rgb_to_gray(r, g, b)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.array_equal(got[5, :width], v)
