"""Colour fills on a 3-byte Rgb8 canvas with the fused rgb_to_gray epilogue,
and fills through the general pipe (Multiply, page soft mask), on the device
through the C-ABI, bit-exact against the oracle (crates/raster/src/pipe/
general.rs:270-344, simd/composite.rs:56-114, crates/pdf_raster/src/
render.rs:1306-1324)."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def smooth_mask(rng, h, w, cells=8):
    """Smooth page-space soft mask: a bilinear upsample of a random lattice."""
    lat = rng.integers(0, 256, (cells + 1, cells + 1)).astype(np.float64)
    ys, xs = np.linspace(0, cells, h), np.linspace(0, cells, w)
    y0, x0 = np.minimum(ys.astype(int), cells - 1), np.minimum(xs.astype(int), cells - 1)
    fy, fx = (ys - y0)[:, None], (xs - x0)[None, :]
    v = (lat[y0][:, x0] * (1 - fy) * (1 - fx) + lat[y0 + 1][:, x0] * fy * (1 - fx)
         + lat[y0][:, x0 + 1] * (1 - fy) * fx + lat[y0 + 1][:, x0 + 1] * fy * fx)
    m = v.astype(np.uint8)
    m[: h // 8] = 255      # bands of mask 255 and 0 too
    m[h // 8: h // 4] = 0
    return m


def paint_attrs(rng, page, general):
    """Random fill flags (even-odd kept from the page; Multiply / soft mask
    when general), colours and opacities for the page's paths."""
    n = page.n_paths
    a = page.attrs.copy()
    flags = a["eo"].astype(np.uint8) & R.FILL_EVEN_ODD
    if general:
        flags |= np.where(rng.random(n) < 0.3, R.FILL_MULTIPLY, 0).astype(np.uint8)
        flags |= np.where(rng.random(n) < 0.3, R.FILL_SOFT_MASK, 0).astype(np.uint8)
    a["eo"] = flags
    a["color"] = rng.integers(0, 256, n)
    a["a_input"] = np.where(rng.random(n) < 0.7, 255, rng.integers(0, 256, n))
    colors = rng.integers(0, 256, (n, 3)).astype(np.uint8)
    return a, colors


def oracle_page(pg, a, colors, flavour, mode, mask, background=255):
    if flavour == R.FLAVOUR_RGB8:
        canv = np.full((pg.height, pg.width, 3), background, np.uint8)
        cols = colors
    else:
        canv = np.full((pg.height, pg.width), background, np.uint8)
        cols = np.repeat(a["color"][:, None], 3, 1)
    _, h = O.fill_page_ex(canv, pg.edges, pg.edge_off, a["eo"], cols, a["a_input"], mode, flavour, mask)
    return canv, h


def run_case(be, pages, attrs, colors, flavour, mode, launch, masks, hashes, gray=False, variant=0):
    rgb = flavour == R.FLAVOUR_RGB8
    db = R.DeviceBatch(be, pages, attrs=np.concatenate(attrs), rgb=rgb,
                       colors=np.concatenate(colors) if rgb else None, gray=gray, soft_masks=masks)
    try:
        db.run(mode, flavour, launch, hashes=hashes, variant=variant)
        got = [db.read_canvas(i) for i in range(len(pages))]
        grays = [db.read_gray(i) for i in range(len(pages))] if gray else None
        gh = db.read_hashes() if hashes else None
    finally:
        db.free()
    for i, pg in enumerate(pages):
        want, wh = oracle_page(pg, attrs[i], colors[i], flavour, mode, masks[i] if masks else None)
        if not np.array_equal(got[i], want):
            bad = np.argwhere(got[i] != want)
            raise AssertionError(f"page {i}: {len(bad)} values differ, first {bad[0]}")
        if gray:
            assert np.array_equal(grays[i], O.rgb_to_gray_page(want)), f"page {i} gray output"
        if hashes:
            off = sum(p.n_paths for p in pages[:i])
            assert np.array_equal(gh[off:off + pg.n_paths], wh), f"page {i} coverage hashes"


@pytest.mark.parametrize("launch", [R.LAUNCH_BATCHED, R.LAUNCH_PLAIN])
@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
def test_colour_fills_on_rgb8_with_gray_output(backend, launch, mode):
    rng = np.random.default_rng(200 + 2 * launch + mode)
    pg = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 5))
    a, col = paint_attrs(rng, pg, general=False)
    run_case(backend, [pg], [a], [col], R.FLAVOUR_RGB8, mode, launch, None, hashes=(mode == 0), gray=True)


@pytest.mark.parametrize("flavour", [R.FLAVOUR_GRAY8, R.FLAVOUR_RGB8_FAST, R.FLAVOUR_RGB8])
@pytest.mark.parametrize("launch", [R.LAUNCH_BATCHED, R.LAUNCH_PLAIN])
def test_general_pipe_fills(backend, flavour, launch):
    # Multiply and soft-masked paths among opaque and translucent AA-path
    # ones, occlusion skipping on (no coverage outputs): a general path is
    # never an occluder, and is itself hidden by a later opaque cover.
    rng = np.random.default_rng(210 + 3 * flavour + launch)
    pg = R.workload(1, 0xC0F10001)[0].subset(range(1, 500, 4))
    a, col = paint_attrs(rng, pg, general=True)
    mask = smooth_mask(rng, pg.height, pg.width)
    run_case(backend, [pg], [a], [col], flavour, R.MODE_DOCUMENTED, launch, [mask], hashes=False,
             gray=(flavour == R.FLAVOUR_RGB8))


def test_general_pipe_with_hashes_and_glyphs(backend):
    rng = np.random.default_rng(220)
    pg = R.workload(2, 0xC0F10002)[0].subset(range(0, 6000, 7))
    a, col = paint_attrs(rng, pg, general=True)
    mask = smooth_mask(rng, pg.height, pg.width)
    run_case(backend, [pg], [a], [col], R.FLAVOUR_RGB8, R.MODE_DOCUMENTED, R.LAUNCH_BATCHED, [mask],
             hashes=True, gray=True, variant=9)


def test_mixed_pages_rgb8_every_tile_gets_gray(backend):
    # Mixed page sizes (tiles at ragged right edges, empty tiles that only the
    # epilogue visits) with colours and the general pipe.
    rng = np.random.default_rng(230)
    pages = [p.subset(range(0, p.n_paths, 6)) for p in R.workload(3, 0xC0F10003, 3)]
    attrs, cols, masks = [], [], []
    for p in pages:
        a, c = paint_attrs(rng, p, general=True)
        attrs.append(a)
        cols.append(c)
        masks.append(smooth_mask(rng, p.height, p.width))
    run_case(backend, pages, attrs, cols, R.FLAVOUR_RGB8, R.MODE_DOCUMENTED, R.LAUNCH_BATCHED, masks,
             hashes=False, gray=True, variant=4)


def test_rgb8_graph_replays(backend):
    rng = np.random.default_rng(240)
    pg = R.workload(1, 0xC0F10001)[0].subset(range(2, 500, 3))
    a, col = paint_attrs(rng, pg, general=True)
    mask = smooth_mask(rng, pg.height, pg.width)
    want, _ = oracle_page(pg, a, col, R.FLAVOUR_RGB8, R.MODE_DOCUMENTED, mask)
    db = R.DeviceBatch(backend, [pg], attrs=a, rgb=True, colors=col, gray=True, soft_masks=[mask])
    g = None
    try:
        g = db.capture(flavour=R.FLAVOUR_RGB8, variant=1)
        for rep in range(3):
            db.run_graph(g)
            assert np.array_equal(db.read_canvas(0), want), f"replay {rep}"
            assert np.array_equal(db.read_gray(0), O.rgb_to_gray_page(want)), f"replay {rep} gray"
            db.run(flavour=R.FLAVOUR_RGB8, launch=R.LAUNCH_PLAIN)  # eager work between replays
    finally:
        if g:
            db.destroy_graph(g)
        db.free()


def test_soft_mask_missing_is_reported(backend):
    pg = R.workload(1, 0xC0F10001)[0].subset(range(0, 40))
    a = pg.attrs.copy()
    a["eo"][3] |= R.FILL_SOFT_MASK
    db = R.DeviceBatch(backend, [pg], attrs=a)
    try:
        for launch in (R.LAUNCH_BATCHED, R.LAUNCH_PLAIN):
            with pytest.raises(R.RrError) as ei:
                db.run(launch=launch)
            assert ei.value.code == R.RR_E_INVARIANT and "soft mask" in str(ei.value)
    finally:
        db.free()


def test_rgb8_needs_colours_and_geometry(backend):
    pg = R.workload(1, 0xC0F10001)[0].subset(range(0, 10))
    db = R.DeviceBatch(backend, [pg])          # 1-byte canvases, no colours
    try:
        backend.begin_page()
        with pytest.raises(R.RrError) as ei:
            backend.record_fill_pages(db.fill_pages(flavour=R.FLAVOUR_RGB8))
        assert ei.value.code == R.RR_E_INVARIANT
        backend.wait_page(backend.submit_page())
    finally:
        db.free()
