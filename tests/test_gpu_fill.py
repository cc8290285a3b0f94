"""Parity of the HIP fill+composite path with the CPU oracle, bit for bit.

Every case renders the same seeded paths through the C-ABI (gfx950 kernels)
and through the oracle, and compares the final u8 canvas and the per-path
coverage hash (sum of a 64-bit mix of (pixel, shape) over every pixel the path
composited with coverage > 0).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_page(page, mode=O.MODE_DOCUMENTED, flavour=O.FLAVOUR_GRAY8, background=255, attrs=None):
    a = page.attrs if attrs is None else attrs
    canvas = np.full((page.height, page.width), background, np.uint8)
    _, hashes = O.fill_page(canvas, page.edges, page.edge_off, a["eo"], a["color"], a["a_input"],
                            mode, flavour)
    return canvas, hashes


def gpu_page(be, page, mode=R.MODE_DOCUMENTED, flavour=R.FLAVOUR_GRAY8, launch=R.LAUNCH_PLAIN,
             background=255, attrs=None):
    dp = R.DevicePage(be, page, attrs)
    try:
        return dp.render(mode, flavour, launch, hashes=True, background=background)
    finally:
        dp.free()


def assert_same(gpu, ora, what):
    gc, gh = gpu
    oc, oh = ora
    if not np.array_equal(gc, oc):
        ys, xs = np.nonzero(gc != oc)
        raise AssertionError(f"{what}: {len(ys)} pixels differ, first at y={ys[0]} x={xs[0]} "
                             f"gpu={gc[ys[0], xs[0]]} oracle={oc[ys[0], xs[0]]}")
    bad = np.nonzero(gh != oh)[0]
    assert len(bad) == 0, f"{what}: coverage hash differs for paths {bad[:10]}"


def hand_page(paths, width=64, height=48):
    """paths: list of (builder_fn, eo, color, a_input)."""
    edges, off, attrs, rows, pts, fls, poff = [], [0], [], [], [], [], [0]
    for build, eo, color, a_in in paths:
        p = O.Path()
        build(p)
        xy, fl = p.arrays()
        e = R.xpath_edges(xy, fl, flatness=0.1)
        edges.append(e.reshape(-1, 4))
        off.append(off[-1] + len(e))
        attrs.append((eo, color, a_in, 0))
        if len(e):
            ys = e[:, [1, 3]]
            rows.append((int(np.floor(ys.min())), int(np.floor(ys.max()))))
        else:
            rows.append((0, -1))
        pts.append(xy.reshape(-1, 2))
        fls.append(fl)
        poff.append(poff[-1] + len(fl))
    return R.Page(width, height, np.concatenate(edges) if edges else np.zeros((0, 4)),
                  np.array(off, np.int64), np.array(attrs, dtype=R.PATH_ATTR_DTYPE),
                  np.array(rows, np.int32).reshape(-1, 2),
                  np.concatenate(pts) if pts else np.zeros((0, 2)),
                  np.concatenate(fls) if fls else np.zeros(0, np.uint8), np.array(poff, np.int64))


def rect(x0, y0, x1, y1):
    def b(p):
        p.move_to(x0, y0)
        p.line_to(x1, y0)
        p.line_to(x1, y1)
        p.line_to(x0, y1)
        p.close(False)
    return b


def donut(p):
    for (a, c) in ((3.3, 40.7), (12.5, 30.2)):
        p.move_to(a, a)
        p.line_to(c, a)
        p.line_to(c, c)
        p.line_to(a, c)
        p.close(True)


def chained_glyph(p):
    p.move_to(20, 10)
    p.curve_to(40, 10, 40, 30, 20, 30)
    p.curve_to(0, 30, 0, 10, 20, 10)
    p.close(False)
    p.move_to(20, 15)
    p.curve_to(30, 15, 30, 25, 20, 25)
    p.curve_to(10, 25, 10, 15, 20, 15)
    p.close(False)


def tri(p):
    p.move_to(0.25, 0.75)
    p.line_to(63.9, 10.1)
    p.line_to(30.5, 47.99)
    p.close(False)


def offpage(p):
    p.move_to(-30, -20)
    p.line_to(-5, -20)
    p.line_to(-10, 60)
    p.close(False)


def edge_touching(p):  # touches the last column and row (Partial clip)
    p.move_to(50.2, 30.1)
    p.line_to(63.999, 35.0)
    p.line_to(60.0, 47.9999)
    p.close(False)


HAND = [
    (rect(1, 1, 5, 5), 0, 0, 255), (rect(7.3, 2.6, 20.9, 13.1), 1, 90, 255),
    (donut, 1, 30, 255), (donut, 0, 200, 128), (chained_glyph, 0, 0, 255),
    (tri, 0, 120, 200), (offpage, 0, 10, 255), (edge_touching, 1, 60, 255),
    (lambda p: (p.move_to(5, 5), p.line_to(6, 5), p.close(False)), 0, 0, 255),  # degenerate
]


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
@pytest.mark.parametrize("flavour", [R.FLAVOUR_GRAY8, R.FLAVOUR_RGB8_FAST])
def test_hand_paths(backend, mode, flavour):
    page = hand_page(HAND)
    assert_same(gpu_page(backend, page, mode, flavour), oracle_page(page, mode, flavour),
                f"hand mode={mode} flavour={flavour}")


def test_each_hand_path_alone(backend):
    for i, item in enumerate(HAND):
        page = hand_page([item])
        assert_same(gpu_page(backend, page), oracle_page(page), f"hand path {i}")


def test_config1_polygons_subset(backend):
    page = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 10))
    assert_same(gpu_page(backend, page), oracle_page(page), "config 1 subset")


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
def test_config1_polygons_literal_vs_documented_subset(backend, mode):
    page = R.workload(1, 0xC0F10001)[0].subset(range(3, 500, 25))
    assert_same(gpu_page(backend, page, mode), oracle_page(page, mode), f"config 1 mode {mode}")


def test_config2_glyphs_subset(backend):
    page = R.workload(2, 0xC0F10002)[0].subset(range(0, 6000, 7))
    assert_same(gpu_page(backend, page), oracle_page(page), "config 2 subset")


def test_config2_close_with_line_variant(backend):
    page = R.workload(102, 0xC0F10002)[0].subset(range(0, 6000, 13))
    assert_same(gpu_page(backend, page), oracle_page(page), "config 2 close-with-line")


def test_config3_page_subset(backend):
    page = R.workload(3, 0xC0F10003, 1)[0]
    page = page.subset(range(0, page.n_paths, 5))
    assert_same(gpu_page(backend, page), oracle_page(page), "config 3 page 0 subset")


def test_random_alpha_and_colors(backend):
    rng = np.random.default_rng(11)
    page = R.workload(1, 0xC0F10001)[0].subset(range(0, 60))
    attrs = page.attrs.copy()
    attrs["a_input"] = rng.integers(0, 256, page.n_paths)
    attrs["color"] = rng.integers(0, 256, page.n_paths)
    for flavour in (R.FLAVOUR_GRAY8, R.FLAVOUR_RGB8_FAST):
        assert_same(gpu_page(backend, page, flavour=flavour, attrs=attrs, background=37),
                    oracle_page(page, flavour=flavour, attrs=attrs, background=37),
                    f"alpha/colour flavour {flavour}")


def test_empty_batch_and_repeat_determinism(backend):
    page = hand_page([])
    c, h = gpu_page(backend, page)
    assert (c == 255).all()
    page = R.workload(1, 0xC0F10001)[0].subset(range(0, 500, 50))
    a = gpu_page(backend, page)
    b = gpu_page(backend, page)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("clip", [(13.5, 7.25, 50.7, 40.2), (0.0, 0.0, 30.0, 47.999), (20.0, 10.0, 20.5, 10.5),
                                  (-5.0, -3.0, 80.0, 90.0)])
@pytest.mark.parametrize("launch", [R.LAUNCH_PLAIN, R.LAUNCH_BATCHED])
def test_clip_rectangles(backend, clip, launch):
    # Clip::new + test_rect + the clip_aa_line clamp on rectangles inside,
    # on the edge of and beyond the page, including one under a pixel wide.
    page = hand_page(HAND)
    want = np.full((page.height, page.width), 255, np.uint8)
    hashes = []
    for i in range(page.n_paths):
        a = page.attrs[i]
        _, h, _ = O.fill(want, page.path_edges(i), a["eo"], a["color"], a["a_input"], clip=np.array(clip))
        hashes.append(h)
    dp = R.DevicePage(backend, page)
    try:
        got = dp.render(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, launch, hashes=True, clip=clip)
    finally:
        dp.free()
    assert_same(got, (want, np.array(hashes, np.uint64)), f"clip {clip} launch {launch}")


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
def test_plain_launch_order_across_streams(backend, mode):
    # Twenty strips in disjoint rows (each starts on its own stream), then
    # half-transparent paths that meet several strips' rows (waits across
    # streams), then strips again over them, then a path meeting everything:
    # the per-pixel composite order is the path order, as on the oracle.
    paths = []
    for k in range(20):
        paths.append((rect(2 + k, 4 * k + 0.3, 40 + k, 4 * k + 3.1), k & 1, (37 * k) & 255, 255))
    for k in range(5):
        paths.append((rect(5 + 3 * k, 10 * k + 0.5, 60 - k, 10 * k + 27.5), 0, (91 * k + 13) & 255, 128))
    for k in range(20):
        paths.append((rect(20 - k * 0.5, 4 * k + 1.2, 50 + k * 0.25, 4 * k + 2.9), 1, (53 * k + 7) & 255, 200))
    paths.append((donut, 0, 30, 100))
    paths.append((rect(0.25, 0.25, 63.5, 79.5), 0, 250, 64))
    page = hand_page(paths, width=64, height=80)
    assert_same(gpu_page(backend, page, mode), oracle_page(page, mode), "strips across streams")


def comb(x0, x1, teeth, y_tip, y_base, y_spine):
    def b(p):
        w = (x1 - x0) / teeth
        p.move_to(x0, y_spine)
        p.line_to(x0, y_base)
        for k in range(teeth):
            p.line_to(x0 + (k + 0.5) * w, y_tip)
            p.line_to(x0 + (k + 1) * w, y_base)
        p.line_to(x1, y_spine)
        p.close(False)
    return b


@pytest.mark.parametrize("mode", [R.MODE_DOCUMENTED, R.MODE_LITERAL])
@pytest.mark.parametrize("variant", [0, 12])
def test_plain_workgroup_shapes(backend, mode, variant):
    # Paths of one band, two and three bands (one workgroup holding every
    # band) and five bands (a workgroup per band): a comb whose tooth rows
    # hold some 300 intersections (those bands go to the one-wave fallback
    # while the spine's band does not), a path wider than the packed 15-bit
    # x range (fallback), a path whose last band holds no drawn row in literal
    # mode, half-transparent overlaps.
    paths = [
        (rect(3.3, 3.3, 120.7, 25.7), 0, 40, 255),                  # 2 bands
        (donut, 1, 90, 200),                                        # 3 bands
        (comb(150, 750, 150, 5.2, 30.4, 45.6), 0, 10, 255),        # 3 bands, fallback in two
        (comb(800, 1400, 150, 20.1, 40.3, 33.2), 1, 170, 128),     # 2 bands, both fall back
        (rect(10.5, 40.2, 8900.5, 57.9), 0, 120, 160),              # 2 bands, wider than 8191 px
        (rect(1500.2, 15.9, 1600.7, 47.05), 0, 200, 255),           # 3 bands, last one empty in literal
        (rect(1700.1, 2.2, 1780.9, 9.8), 1, 60, 255),               # 1 band
        (rect(1800.4, 2.6, 1900.2, 77.4), 0, 220, 100),             # 5 bands
        (chained_glyph, 0, 0, 255),
    ]
    page = hand_page(paths, width=9000, height=80)
    want = oracle_page(page, mode)
    db = R.DeviceBatch(backend, [page])
    try:
        db.run(mode, R.FLAVOUR_GRAY8, R.LAUNCH_PLAIN, hashes=True, variant=variant)
        got = (db.read_canvas(0), db.read_hashes())
    finally:
        db.free()
    assert_same(got, want, f"workgroup shapes, variant {variant}")


_FEW_QUEUES = r"""
import sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/pdf-raster_amd"]
import rr_hip as R
from oracle import oracle as O
pg = R.workload(2, 0xC0F10002)[0].subset(range(0, 6000, 7))
be = R.Backend(0)
dp = R.DevicePage(be, pg)
got, gh = dp.render(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, R.LAUNCH_PLAIN, hashes=True)
dp.free()
be.close()
want = np.full((pg.height, pg.width), 255, np.uint8)
_, wh = O.fill_page(want, pg.edges, pg.edge_off, pg.attrs["eo"], pg.attrs["color"], pg.attrs["a_input"],
                    O.MODE_DOCUMENTED, O.FLAVOUR_GRAY8)
assert np.array_equal(got, want), f"{int((got != want).sum())} pixels differ"
assert np.array_equal(gh, np.asarray(wh, np.uint64)), "coverage hashes differ"
print("ok", pg.n_paths)
"""


def test_plain_route_with_ten_hardware_queues():
    # A process started with GPU_MAX_HW_QUEUES=10 spreads the plain launches
    # over 8 side streams (two fewer than the queues): 858 glyphs of the glyph
    # page against the oracle in a fresh process.
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, GPU_MAX_HW_QUEUES="10")
    r = subprocess.run([sys.executable, "-c", _FEW_QUEUES, root], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    assert r.stdout.startswith("ok")
