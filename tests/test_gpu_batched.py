"""The batched persistent launch gives the same pixels and per-path coverage
hashes as the oracle (and hence as the one-path-per-launch route)."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_pages(pages, mode=O.MODE_DOCUMENTED, flavour=O.FLAVOUR_GRAY8):
    canv, hs = [], []
    for pg in pages:
        c = np.full((pg.height, pg.width), 255, np.uint8)
        a = pg.attrs
        _, h = O.fill_page(c, pg.edges, pg.edge_off, a["eo"], a["color"], a["a_input"], mode, flavour)
        canv.append(c)
        hs.append(h)
    return canv, np.concatenate(hs)


def check(be, pages, mode=R.MODE_DOCUMENTED, flavour=R.FLAVOUR_GRAY8, launch=R.LAUNCH_BATCHED, hashes=True,
          variant=0):
    db = R.DeviceBatch(be, pages)
    try:
        db.run(mode, flavour, launch, hashes=hashes, variant=variant)
        got = [db.read_canvas(i) for i in range(len(pages))]
        gh = db.read_hashes()
    finally:
        db.free()
    oc, oh = oracle_pages(pages, mode, flavour)
    for i, (g, o) in enumerate(zip(got, oc)):
        if not np.array_equal(g, o):
            ys, xs = np.nonzero(g != o)
            raise AssertionError(f"page {i}: {len(ys)} px differ, first y={ys[0]} x={xs[0]}")
    if hashes:
        bad = np.nonzero(gh != oh)[0]
        assert len(bad) == 0, f"coverage hashes differ for paths {bad[:10]}"


def test_batched_polygon_subset(backend):
    check(backend, [R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 7))])


def test_batched_glyph_subset(backend):
    check(backend, [R.workload(2, 0xC0F10002)[0].subset(range(0, 6000, 5))])


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
def test_batched_mixed_pages(backend, mode):
    pages = R.workload(3, 0xC0F10003, 3)
    pages = [pg.subset(range(0, pg.n_paths, 4)) for pg in pages]
    check(backend, pages, mode)


def test_batched_rgb8_flavour(backend):
    pages = [R.workload(1, 0xC0F10001)[0].subset(range(1, 500, 11))]
    check(backend, pages, flavour=R.FLAVOUR_RGB8_FAST)


def test_plain_multi_page_matches(backend):
    pages = R.workload(3, 0xC0F10003, 2)
    pages = [pg.subset(range(0, pg.n_paths, 9)) for pg in pages]
    check(backend, pages, launch=R.LAUNCH_PLAIN)


@pytest.mark.parametrize("variant", range(len(R.variants())))
def test_every_variant_is_bit_exact(backend, variant):
    pages = R.workload(3, 0xC0F10003, 2)
    pages = [pg.subset(range(0, pg.n_paths, 6)) for pg in pages]
    db = R.DeviceBatch(backend, pages)
    try:
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=variant)
        got = [db.read_canvas(i) for i in range(len(pages))]
        gh = db.read_hashes()
        db.run(launch=R.LAUNCH_PLAIN, hashes=True, variant=variant)
        got_p = [db.read_canvas(i) for i in range(len(pages))]
        gh_p = db.read_hashes()
    finally:
        db.free()
    oc, oh = oracle_pages(pages)
    for g, gp, o in zip(got, got_p, oc):
        assert np.array_equal(g, o) and np.array_equal(gp, o)
    assert np.array_equal(gh, oh) and np.array_equal(gh_p, oh)


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
def test_batched_without_coverage_outputs(backend, mode):
    # Without per-path coverage outputs the composite pass starts each tile at
    # the last opaque path covering it fully; the pixels must not change.
    # Whole pages, so that large opaque paths overlap many others.
    pages = R.workload(3, 0xC0F10003, 3)
    oc, _ = oracle_pages(pages, mode)
    be = R.Backend(0)
    db = R.DeviceBatch(be, pages)
    try:
        for variant in (0, 2, 7, 11):
            db.run(mode, R.FLAVOUR_GRAY8, R.LAUNCH_BATCHED, variant=variant)
            for i, o in enumerate(oc):
                g = db.read_canvas(i)
                if not np.array_equal(g, o):
                    ys, xs = np.nonzero(g != o)
                    raise AssertionError(f"variant {variant} page {i}: {len(ys)} px differ, first y={ys[0]} x={xs[0]}")
    finally:
        db.free()
        be.close()


def test_interval_pool_grow_and_redo():
    # A fresh context sizes its interval pool at 24 per (path, band); whole
    # mixed pages with coverage hashes (no occlusion skipping) need more, so
    # the first recording grows the pool and redoes the span pass.
    pages = R.workload(3, 0xC0F10003, 3)
    oc, oh = oracle_pages(pages)
    be = R.Backend(0)
    db = R.DeviceBatch(be, pages)
    try:
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, counts=True, variant=2)
        assert be.last_batch_stats()[1]["reruns"] >= 1
        for i, o in enumerate(oc):
            assert np.array_equal(db.read_canvas(i), o), f"page {i}"
        assert np.array_equal(db.read_hashes(), oh)
    finally:
        db.free()
        be.close()


def edge_page(width, height, paths):
    """A page from raw add_segment tuples: paths = [(edges (n, 4) in device
    pixels, eo, color, a_input)]."""
    edges = [np.asarray(e, np.float64).reshape(-1, 4) for e, *_ in paths]
    off = np.zeros(len(paths) + 1, np.int64)
    off[1:] = np.cumsum([len(e) for e in edges])
    attrs = np.array([(eo, color, a_in, 0) for _, eo, color, a_in in paths], dtype=R.PATH_ATTR_DTYPE)
    rows = np.array([(int(np.floor(e[:, [1, 3]].min())), int(np.floor(e[:, [1, 3]].max()))) for e in edges],
                    np.int32).reshape(-1, 2)
    return R.Page(width, height, np.concatenate(edges), off, attrs, rows, np.zeros((0, 2)),
                  np.zeros(0, np.uint8), np.zeros(len(paths) + 1, np.int64))


def rect_edges(x0, y0, x1, y1):
    return [[x0, y0, x1, y0], [x1, y0, x1, y1], [x1, y1, x0, y1], [x0, y1, x0, y0]]


def canvases_match(db, pages, oc, what):
    for i, o in enumerate(oc):
        g = db.read_canvas(i)
        if not np.array_equal(g, o):
            ys, xs = np.nonzero(g != o)
            raise AssertionError(f"{what} page {i}: {len(ys)} px differ, first y={ys[0]} x={xs[0]}")


def test_unbalanced_rows_with_occlusion(backend):
    # A path of one edge leaves every row it crosses unbalanced: its span runs
    # to the canvas edge and the path's columns are widened.  Opaque
    # rectangles drawn later cover whole composite tiles over such paths'
    # own columns, so with occlusion on (no coverage outputs) some of their
    # bands can be skipped before another band widens the path; the fixup
    # pass must redo those.  Pixels must match the oracle in every case.
    rng = np.random.default_rng(7)
    pages = []
    for _ in range(4):
        paths = []
        for i in range(24):
            x = rng.uniform(2, 250)
            paths.append(([x, rng.uniform(0, 20), x + rng.uniform(-3, 3), rng.uniform(40, 96)], i % 2,
                          int(rng.integers(0, 256)), 255))
        for _ in range(24):
            b, t = int(rng.integers(0, 6)), int(rng.integers(0, 3))
            paths.append((rect_edges(64 * t - 0.5, 16 * b - 0.5, 64 * t + 128.5, 16 * b + 16.5), 0,
                          int(rng.integers(0, 256)), 255))
        pages.append(edge_page(256, 96, paths))
    oc, oh = oracle_pages(pages)
    db = R.DeviceBatch(backend, pages)
    try:
        for variant in (0, 4, 6):
            db.run(launch=R.LAUNCH_BATCHED, variant=variant)
            canvases_match(db, pages, oc, f"variant {variant}")
        db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=4)
        canvases_match(db, pages, oc, "hashes")
        assert np.array_equal(db.read_hashes(), oh)
    finally:
        db.free()


def comb_edges(x0, y0, y1, teeth, dx):
    # a horizontal zigzag outline: every row crosses about 2 * teeth edges
    # whose x ranges do not touch, so rows hold that many entries
    pts = [(x0 + k * dx, y0 if k % 2 else y1) for k in range(2 * teeth + 1)]
    pts += [(pts[-1][0], y1 + 4), (x0, y1 + 4)]
    return [[*pts[i], *pts[(i + 1) % len(pts)]] for i in range(len(pts))]


def test_rows_beyond_the_lds_slots_go_to_the_wide_pass(backend):
    # Rows with more state-changing entries than the main span pass's LDS
    # slots are deferred to the wide pass (32 slots, then global memory).
    rng = np.random.default_rng(11)
    pages = []
    for _ in range(2):
        paths = []
        for i in range(12):
            teeth = int(rng.integers(3, 40))
            x0, y0 = rng.uniform(0, 40), rng.uniform(0, 60)
            paths.append((comb_edges(x0, y0, y0 + rng.uniform(10, 60), teeth, rng.uniform(2, 5)), i % 3 == 0,
                          int(rng.integers(0, 256)), 255))
        pages.append(edge_page(320, 140, paths))
    oc, oh = oracle_pages(pages)
    db = R.DeviceBatch(backend, pages)
    try:
        for variant in (0, 5):  # 4 and 8 LDS slots
            db.run(launch=R.LAUNCH_BATCHED, hashes=True, variant=variant)
            assert backend.last_batch_stats()[1]["deferred"] > 0
            canvases_match(db, pages, oc, f"variant {variant}")
            assert np.array_equal(db.read_hashes(), oh)
            db.run(launch=R.LAUNCH_BATCHED, variant=variant)
            canvases_match(db, pages, oc, f"variant {variant} without coverage outputs")
    finally:
        db.free()


def extreme_page():
    """Edges far outside the int32 and 16.16 ranges, nearly horizontal slopes,
    and NaN / infinite coordinates (the f64 -> int conversions of the edge
    setup saturate, and NaN edges are dropped, as in the reference)."""
    inf, nan = np.inf, np.nan
    W, H = 96, 80
    paths = [
        [(10, 10, 1e12, 40), (1e12, 40, 12, 70), (12, 70, 10, 10)],
        [(-3e9, 5, 50, 5.5), (50, 5.5, 60, 30), (60, 30, -3e9, 29.5), (-3e9, 29.5, -3e9, 5)],
        [(5, 50, 90, 50 + 1e-9), (90, 50 + 1e-9, 90, 60), (90, 60, 5, 60 - 1e-9), (5, 60 - 1e-9, 5, 50)],
        [(20, -1e15, 30, 1e15), (30, 1e15, 40, 1e15), (40, 1e15, 35, -1e15), (35, -1e15, 20, -1e15)],
        [(60, 60, 80, 62), (80, 62, nan, 70), (nan, 70, 70, 75), (70, 75, inf, 78), (inf, 78, 60, 60),
         (60, 60, 60, -inf), (65, 65, 66, 66)],
        [(-1e10, 2, -1e10, 70), (-1e10, 70, 8, 72), (8, 72, 8, 2), (8, 2, -1e10, 2)],
        [(536870911.75, 1, 2, 3), (2, 3, 3, 9), (3, 9, 536870912.25, 1)],
    ]
    edges = np.array([e for p in paths for e in p], np.float64)
    off = np.cumsum([0] + [len(p) for p in paths]).astype(np.int64)
    n = len(paths)
    attrs = np.array([(i % 2, (40 * i) % 256, 255 if i % 3 else 170, 0) for i in range(n)],
                     dtype=R.PATH_ATTR_DTYPE)
    rows = np.tile(np.array([[0, H - 1]], np.int32), (n, 1))
    return R.Page(W, H, edges, off, attrs, rows, np.zeros((0, 2)), np.zeros(0, np.uint8),
                  np.zeros(n + 1, np.int64))


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
@pytest.mark.parametrize("launch", [R.LAUNCH_BATCHED, R.LAUNCH_PLAIN])
def test_extreme_and_non_finite_coordinates(backend, mode, launch):
    check(backend, [extreme_page()], mode, launch=launch)


def test_empty_pages_and_pathless_paths_in_a_batch(backend):
    # A page without paths between two others, a path without edges, and a
    # page whose width is not a multiple of the composite tile.
    poly = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 25))
    glyph = R.workload(2, 0xC0F10002, max_pages=1)[0].subset(range(0, 6000, 300))
    empty = R.Page(333, 77, np.zeros((0, 4)), np.zeros(1, np.int64), np.zeros(0, dtype=R.PATH_ATTR_DTYPE),
                   np.zeros((0, 2), np.int32), np.zeros((0, 2)), np.zeros(0, np.uint8), np.zeros(1, np.int64))
    e = glyph.edges.reshape(-1, 4)
    off = np.concatenate([glyph.edge_off[:3], glyph.edge_off[2:]])  # path 2 has no edges
    attrs = np.concatenate([glyph.attrs[:2], glyph.attrs[1:2], glyph.attrs[2:]])
    rows = np.concatenate([glyph.host_rows[:2], np.array([[0, -1]], np.int32), glyph.host_rows[2:]])
    holed = R.Page(glyph.width - 37, glyph.height, e, off, attrs, rows, np.zeros((0, 2)), np.zeros(0, np.uint8),
                   np.zeros(len(off), np.int64))
    check(backend, [poly, empty, holed])
