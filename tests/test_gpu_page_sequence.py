"""One device page buffer across fills, image placement and more fills,
recorded as one page and downloaded once: the canvas equals the oracle running
the same operations in the same order (fill/mod.rs fills, blit_image's general
affine path, page/mod.rs:1558-1975)."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_fills_images_fills_on_one_resident_page(backend):
    poly = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 5))                # 100 polygons
    glyph = R.workload(2, 0xC0F10002, max_pages=1)[0].subset(range(0, 6000, 20))  # 300 glyphs
    iset = R.images(11, 2, 2)                                                      # a CMYK and an RGB image
    assert (poly.width, poly.height) == (glyph.width, glyph.height) == (iset.page_w, iset.page_h)
    W, H = poly.width, poly.height
    stride = R.stride_for(W)
    be = backend
    canvas = be.to_device(np.full((H, stride), 255, np.uint8))
    dp = R.DevicePage(be, poly)
    db = R.DeviceBatch(be, [glyph])
    di = R.DeviceImages(be, iset)
    try:
        be.begin_page()
        b1 = dp.batch(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, R.LAUNCH_PLAIN)
        b1.canvas = canvas
        be.record_fill_paths(b1)
        di.pages[0].canvas = canvas
        be.record_place_images(di.batch())
        db.descs[0].canvas = canvas
        be.record_fill_pages(db.fill_pages(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, R.LAUNCH_BATCHED))
        be.wait_page(be.submit_page())
        got = be.download(canvas, np.zeros((H, stride), np.uint8))[:, :W]
    finally:
        dp.free()
        db.free()
        di.free()
        be.free(canvas)
    want = np.full((H, W), 255, np.uint8)
    for pg in (poly,):
        a = pg.attrs
        O.fill_page(want, pg.edges, pg.edge_off, a["eo"], a["color"], a["a_input"], O.MODE_DOCUMENTED,
                    O.FLAVOUR_GRAY8)
    for im in iset.images:
        grid, g = (iset.grid4, iset.grid4_n) if im.channels == 4 else (iset.grid3, iset.grid3_n)
        assert O.place_image(want, im.pixels, im.smask, im.ctm, im.blend, grid, g) >= 0
    a = glyph.attrs
    O.fill_page(want, glyph.edges, glyph.edge_off, a["eo"], a["color"], a["a_input"], O.MODE_DOCUMENTED,
                O.FLAVOUR_GRAY8)
    if not np.array_equal(got, want):
        ys, xs = np.nonzero(got != want)
        raise AssertionError(f"{len(ys)} px differ, first y={ys[0]} x={xs[0]}")
