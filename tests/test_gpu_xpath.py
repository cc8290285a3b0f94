"""Device path ingestion (XPath::new + flatten_curve on the GPU) through the
C-ABI: edge lists bit-identical to the oracle's XPath restatement, and the
fill of device-ingested edges identical to the oracle's fill."""
import numpy as np
import pytest

import rr_hip as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def device_xpath(be, pts, flags, pt_off, matrices=None, flatness=0.1, edge_cap=None):
    """Run rr_hip_xpath_edges over host arrays; returns (rc, n_edges, edges, edge_off)."""
    n_paths, n_pts = len(pt_off) - 1, len(flags)
    d_xy = be.to_device(np.ascontiguousarray(pts, np.float64).reshape(-1))
    d_fl = be.to_device(np.ascontiguousarray(flags, np.uint8))
    d_off = be.to_device(np.ascontiguousarray(pt_off, np.int64))
    d_m = be.to_device(np.ascontiguousarray(matrices, np.float64).reshape(-1)) if matrices is not None else None
    d_eoff = be.alloc(8 * (n_paths + 1))
    bufs = [d_xy, d_fl, d_off, d_eoff] + ([d_m] if d_m else [])
    try:
        if edge_cap is None:
            rc, total = be.xpath_edges(d_xy, d_fl, d_off, n_paths, n_pts, None, 0, d_eoff, d_m, flatness,
                                       check=False)
            if rc not in (0, R.RR_E_CAPACITY):
                return rc, total, None, None
            edge_cap = total
        d_e = be.alloc(max(32 * edge_cap, 32))
        bufs.append(d_e)
        rc, total = be.xpath_edges(d_xy, d_fl, d_off, n_paths, n_pts, d_e, edge_cap, d_eoff, d_m, flatness,
                                   check=False)
        eoff = be.download(d_eoff, np.zeros(n_paths + 1, np.int64))
        edges = be.download(d_e, np.zeros((max(edge_cap, 1), 4), np.float64))[:min(total, edge_cap)]
        return rc, total, edges, eoff
    finally:
        for b in bufs:
            be.free(b)


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                                                 np.ascontiguousarray(b).view(np.uint64))


@pytest.mark.parametrize("cfg,seed", [(1, 0xC0F10001), (2, 0xC0F10002), (3, 0xC0F10003)])
def test_workload_pages_match_host_and_oracle(backend, cfg, seed):
    page = R.workload(cfg, seed, max_pages=1)[0]
    rc, total, edges, eoff = device_xpath(backend, page.pts, page.pt_flags, page.pt_off)
    assert rc == 0 and total == page.n_edges
    assert np.array_equal(eoff, page.edge_off)
    assert same_bits(edges, page.edges.reshape(-1, 4))
    # the host edges are the oracle's; spot-check the device ones against it directly
    for i in range(0, page.n_paths, max(1, page.n_paths // 25)):
        xy, fl = page.path_points(i)
        assert same_bits(edges[eoff[i]:eoff[i + 1]], O.xpath_edges(xy, fl, flatness=0.1)), f"path {i}"


def random_paths(rng, n_paths):
    """PathBuilder outputs plus hand-made flag sequences the walk must treat
    exactly as the reference: points before the first subpath, subpaths ending
    in curves (which run on into the next point), truncated curves, lone
    move-tos, repeated closes."""
    pts, flags, off = [], [], [0]
    for k in range(n_paths):
        if k % 5 == 4:
            n = int(rng.integers(1, 12))
            xy = rng.uniform(-50, 150, (n, 2))
            fl = rng.choice(np.array([0, 1, 2, 3, 8, 9, 10, 4, 5, 6, 12], np.uint8), n)
        else:
            p = O.Path()
            for _ in range(int(rng.integers(1, 4))):
                p.move_to(*rng.uniform(0, 100, 2))
                for _ in range(int(rng.integers(0, 7))):
                    if rng.random() < 0.5:
                        p.line_to(*rng.uniform(0, 100, 2))
                    else:
                        p.curve_to(*rng.uniform(-20, 120, 6))
                if rng.random() < 0.6:
                    p.close(bool(rng.random() < 0.5))
            xy, fl = p.arrays()
            xy = xy.reshape(-1, 2)
        pts.append(xy)
        flags.append(fl)
        off.append(off[-1] + len(fl))
    mats = np.stack([np.array([rng.uniform(0.2, 3), rng.uniform(-1, 1), rng.uniform(-1, 1),
                               rng.uniform(0.2, 3), rng.uniform(-40, 40), rng.uniform(-40, 40)])
                     for _ in range(n_paths)])
    return np.concatenate(pts), np.concatenate(flags), np.array(off, np.int64), mats


@pytest.mark.parametrize("seed,flatness", [(1, 0.1), (2, 0.5), (3, 0.01), (4, 0.0)])
def test_random_paths_with_matrices(backend, seed, flatness):
    rng = np.random.default_rng(seed)
    pts, flags, off, mats = random_paths(rng, 300)
    rc, total, edges, eoff = device_xpath(backend, pts, flags, off, mats, flatness)
    assert rc == 0
    for i in range(len(off) - 1):
        want = O.xpath_edges(pts[off[i]:off[i + 1]], flags[off[i]:off[i + 1]], mats[i], flatness=flatness)
        assert same_bits(edges[eoff[i]:eoff[i + 1]], want.reshape(-1, 4)), f"seed {seed} path {i}"
    assert total == eoff[-1]


def test_long_malformed_flag_runs(backend):
    """Flag sequences of 65 to 400 points: the walk's state (inside a
    subpath, or skipping a curve's control points) and the subpath start are
    carried across 64-point chunks, and a curve cut short at a chunk's first
    point closes the subpath at the previous chunk's last point."""
    rng = np.random.default_rng(23)
    choices = np.array([0, 0, 0, 8, 8, 8, 1, 2, 3, 9, 10, 4, 5, 6, 12], np.uint8)
    pts, flags, off = [], [], [0]
    for k in range(60):
        n = int(rng.integers(65, 400))
        fl = rng.choice(choices, n)
        if k % 3 == 0:
            fl[:] = 8  # all curve flags: skips running over chunk boundaries
            fl[0] = 1
        if k % 3 == 1:
            fl[-(k % 5 + 1)] = 8  # a curve cut short near the end
        pts.append(rng.uniform(-50, 150, (n, 2)))
        flags.append(fl)
        off.append(off[-1] + n)
    pts, flags, off = np.concatenate(pts), np.concatenate(flags), np.array(off, np.int64)
    rc, total, edges, eoff = device_xpath(backend, pts, flags, off)
    assert rc == 0 and total == eoff[-1]
    for i in range(len(off) - 1):
        want = O.xpath_edges(pts[off[i]:off[i + 1]], flags[off[i]:off[i + 1]], flatness=0.1)
        assert same_bits(edges[eoff[i]:eoff[i + 1]], want.reshape(-1, 4)), f"path {i}"


def test_curves_beyond_half_range(backend):
    """Curves with a control point beyond f64::MAX/2, infinite or NaN take
    f64::midpoint's general form (and, never flat, split down to single
    slots); f64::MAX/2 itself and in-range curves of the same batch take
    (a + b) / 2.  Bit-identical to the oracle either way, except that a NaN
    the arithmetic creates (inf/2 - inf/2) carries the processor's default
    NaN bits (x86 sets the sign, the GPU does not), so NaNs compare by
    position."""
    special = [1.5e308, -1.7e308, 9e307, np.inf, np.nan, 8.98846567431157953865e307, -8.98846567431157953865e307]
    rng = np.random.default_rng(11)
    pts, flags, off = [], [], [0]
    for k in range(42):
        p = O.Path()
        p.move_to(*rng.uniform(0, 100, 2))
        c = rng.uniform(-50, 150, 6)
        if k % 3 != 2:
            c[k % 6] = special[k % len(special)]
        p.curve_to(*c)
        p.line_to(*rng.uniform(0, 100, 2))
        p.close(False)
        xy, fl = p.arrays()
        pts.append(xy.reshape(-1, 2))
        flags.append(fl)
        off.append(off[-1] + len(fl))
    pts, flags, off = np.concatenate(pts), np.concatenate(flags), np.array(off, np.int64)
    rc, total, edges, eoff = device_xpath(backend, pts, flags, off)
    assert rc == 0 and total == eoff[-1]
    for i in range(len(off) - 1):
        want = O.xpath_edges(pts[off[i]:off[i + 1]], flags[off[i]:off[i + 1]], flatness=0.1).reshape(-1, 4)
        got = edges[eoff[i]:eoff[i + 1]]
        assert got.shape == want.shape, f"path {i}"
        assert np.array_equal(np.isnan(got), np.isnan(want)), f"path {i}"
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok].view(np.uint64), want[ok].view(np.uint64)), f"path {i}"


def test_capacity_empty_and_bad_offsets(backend):
    rng = np.random.default_rng(9)
    pts, flags, off, _ = random_paths(rng, 20)
    rc, total, _, _ = device_xpath(backend, pts, flags, off)
    assert rc == 0 and total > 1
    rc, total2, edges, eoff = device_xpath(backend, pts, flags, off, edge_cap=total - 1)
    assert rc == R.RR_E_CAPACITY and total2 == total and eoff[-1] == total
    rc, total, _, eoff = device_xpath(backend, np.zeros((0, 2)), np.zeros(0, np.uint8), np.zeros(1, np.int64))
    assert rc == 0 and total == 0 and eoff[0] == 0
    bad = off.copy()
    bad[3], bad[4] = bad[4], bad[3]
    rc, _, _, _ = device_xpath(backend, pts, flags, bad)
    assert rc == R.RR_E_INVARIANT


def test_fill_of_device_ingested_glyphs(backend):
    """Device ingestion feeding the batched fill: the canvas and coverage
    hashes equal the oracle's fill of the oracle's edges."""
    page = R.workload(2, 0xC0F10002, max_pages=1)[0].subset(range(0, 6000, 7))
    rc, total, edges, eoff = device_xpath(backend, page.pts, page.pt_flags, page.pt_off)
    assert rc == 0
    ingested = R.Page(page.width, page.height, edges, eoff, page.attrs, page.host_rows, page.pts,
                      page.pt_flags, page.pt_off)
    db = R.DeviceBatch(backend, [ingested])
    try:
        db.run(R.MODE_DOCUMENTED, R.FLAVOUR_GRAY8, R.LAUNCH_BATCHED, hashes=True)
        got, gh = db.read_canvas(0), db.read_hashes()
    finally:
        db.free()
    canvas = np.full((page.height, page.width), 255, np.uint8)
    a = page.attrs
    _, oh = O.fill_page(canvas, page.edges, page.edge_off, a["eo"], a["color"], a["a_input"],
                        O.MODE_DOCUMENTED, O.FLAVOUR_GRAY8)
    assert np.array_equal(got, canvas)
    assert np.array_equal(gh, oh)
