"""Pins the oracle's anti-aliased fill path -- XPathScanner construction,
render_aa_line, the fill_impl AA loop in both row modes and draw_aa_line --
with checks that do not restate its code:

* closed forms: rectangles with corners on the quarter-pixel grid, whose
  coverage follows from the reference's rules by hand (below), in both AA row
  modes, hairline rectangles included;
* invariants: a closed path's counts sum to zero on every scanner row, rows
  are sorted by x0, non-zero and even-odd agree on convex polygons;
* a brute-force sampler: away from the columns an edge touches in an AA row,
  a cell is covered exactly when the winding number of the point (c + 0.5, r)
  -- segments with y_min < r <= y_max, left of the point -- is inside;
* fp_from_f64's truncation toward zero (Rust's `as i32`: truncating,
  saturating, NaN -> 0).

Then the mutation check: six builds of the oracle, each with one planted
mistake (RRO_MUTANT in oracle/rr_oracle.c), must each fail a named check.

Rectangle closed form (AA space, corners X0 < X1, Y0 < Y1 integers), from
crates/raster/src/scanner/mod.rs:511-642 and 393-467.  A vertical edge gets
its count on rows r with y_min < r, so row Y0 holds only count-0 entries and
draws nothing.  Rows Y0 < r < Y1 hold [X0, X0] and [X1, X1] with opposite
counts; the walk opens at X0 and closes at the x1 + 1 of the closing entry,
so AA columns X0..X1 (inclusive) are covered -- unless X1 - X0 <= 1, where
the two entries touch, merge to count 0 and nothing is covered.  Row Y1
depends on the order the edges are written (write_intersect merges into the
row's last entry only): the bottom edge's count-0 entry [X0, X1] touches
both vertical entries, so written between or before them it merges them into
one entry of count 0 and the row draws nothing; written after both, it
merges into the second only, the row keeps [X0, X0] and [X0, X1] with
opposite counts and is covered like the rows above it.  Output pixel (x, y)
then gets count = sum over its AA rows of |[X0, X1 + 1) & [4x, 4x + 4)| and
shape AA_GAMMA[count]; its AA rows are 4y..4y+3 (documented) or, in the
literal reading of fill/mod.rs:176-203, 16y+12..16y+15, drawn only when the AA
row 4y + 3 lies in the scanner's rows [Y0, Y1].
"""
from fractions import Fraction
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

GAMMA = np.array([round((i / 16) ** 1.5 * 255) for i in range(17)], np.uint8)


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------

def rect_edges(X0, Y0, X1, Y1, rng):
    """A closed rectangle (AA corners) as add_segment tuples in pixel space,
    random orientation and starting corner; also whether the bottom edge is
    written after both vertical edges."""
    c = [(X0, Y0), (X1, Y0), (X1, Y1), (X0, Y1)]
    if rng.random() < 0.5:
        c = c[::-1]
    k = int(rng.integers(0, 4))
    c = c[k:] + c[:k]
    e = np.array([[c[i][0] / 4, c[i][1] / 4, c[(i + 1) % 4][0] / 4, c[(i + 1) % 4][1] / 4]
                  for i in range(4)], np.float64)
    bottom = [i for i in range(4) if e[i, 1] == e[i, 3] == Y1 / 4][0]
    verticals = [i for i in range(4) if e[i, 0] == e[i, 2]]
    return e, bottom > max(verticals)


def rect_shape(W, H, X0, Y0, X1, Y1, mode, bottom_last):
    rows = set(range(Y0 + 1, Y1)) if X1 - X0 >= 2 else set()
    if bottom_last and X1 - X0 >= 2:
        rows.add(Y1)
    shape = np.zeros((H, W), np.uint8)
    for y in range(H):
        if mode == O.MODE_LITERAL:
            if not Y0 <= 4 * y + 3 <= Y1:
                continue
            aa = range(16 * y + 12, 16 * y + 16)
        else:
            aa = range(4 * y, 4 * y + 4)
        n = sum(1 for r in aa if r in rows)
        if n == 0:
            continue
        for x in range(W):
            ov = max(0, min(X1 + 1, 4 * x + 4) - max(X0, 4 * x))
            if ov:
                shape[y, x] = GAMMA[n * ov]
    return shape


def check_rectangles(mode, hairline=False, n=60, seed=1):
    rng = np.random.default_rng(seed + 10 * mode + 100 * hairline)
    W, H = 23, 19 if mode == O.MODE_DOCUMENTED else 80
    for _ in range(n):
        if hairline:
            X0 = int(rng.integers(1, 4 * W - 8))
            X1 = X0 + int(rng.integers(0, 2))
        else:
            X0 = int(rng.integers(1, 4 * W - 12))
            X1 = int(rng.integers(X0 + 2, 4 * W - 5))
        ylim = 4 * H - 5 if mode == O.MODE_DOCUMENTED else 4 * (H // 4) - 5
        Y0 = int(rng.integers(1, ylim - 4))
        Y1 = int(rng.integers(Y0 + 2, ylim))
        edges, bottom_last = rect_edges(X0, Y0, X1, Y1, rng)
        canvas = np.full((H, W), 255, np.uint8)
        _, _, shape = O.fill(canvas, edges, int(rng.integers(0, 2)), 0, 255, mode=mode, want_shape=True)
        want = rect_shape(W, H, X0, Y0, X1, Y1, mode, bottom_last)
        if not np.array_equal(shape, want):
            ys, xs = np.nonzero(shape != want)
            raise AssertionError(f"rect AA [{X0},{X1}]x[{Y0},{Y1}] mode {mode}: {len(ys)} px differ, "
                                 f"first ({xs[0]},{ys[0]}) got {shape[ys[0], xs[0]]} want {want[ys[0], xs[0]]}")


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------

def random_polygon(rng, k, W, H):
    pts = np.stack([rng.uniform(-3, W + 3, k), rng.uniform(-3, H + 3, k)], 1)
    return np.array([[*pts[i], *pts[(i + 1) % k]] for i in range(k)], np.float64)


def check_row_counts_sum_to_zero(n=80, seed=2):
    rng = np.random.default_rng(seed)
    W, H = 40, 30
    for _ in range(n):
        e = random_polygon(rng, int(rng.integers(3, 12)), W, H)
        s = O.Scanner(e, int(rng.integers(0, 2)), 0, 4 * H - 1, aa=True)
        _, _, y0, y1 = s.bbox()
        for r in range(y0, y1 + 1):
            row = s.row(r)
            tot = sum(c for _, _, c in row) & 0xffffffff
            assert tot == 0, f"row {r}: counts sum to {tot}"


def check_rows_sorted(n=80, seed=3):
    rng = np.random.default_rng(seed)
    W, H = 40, 30
    for _ in range(n):
        e = random_polygon(rng, int(rng.integers(3, 16)), W, H)
        s = O.Scanner(e, 0, 0, 4 * H - 1, aa=True)
        _, _, y0, y1 = s.bbox()
        for r in range(y0, y1 + 1):
            x0s = [a for a, _, _ in s.row(r)]
            assert x0s == sorted(x0s), f"row {r} not sorted by x0: {x0s}"


def check_nz_equals_eo_on_convex(n=60, seed=4):
    rng = np.random.default_rng(seed)
    W, H = 37, 29
    for _ in range(n):
        k = int(rng.integers(3, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        cx, cy, rx, ry = rng.uniform(5, W - 5), rng.uniform(5, H - 5), rng.uniform(1, 15), rng.uniform(1, 15)
        pts = np.stack([cx + rx * np.cos(ang), cy + ry * np.sin(ang)], 1)
        e = np.array([[*pts[i], *pts[(i + 1) % k]] for i in range(k)], np.float64)
        for mode in (O.MODE_DOCUMENTED, O.MODE_LITERAL):
            _, _, s_nz = O.fill(np.full((H, W), 255, np.uint8), e, 0, 0, mode=mode, want_shape=True)
            _, _, s_eo = O.fill(np.full((H, W), 255, np.uint8), e, 1, 0, mode=mode, want_shape=True)
            assert np.array_equal(s_nz, s_eo), "non-zero and even-odd differ on a convex polygon"


def check_fp_truncates():
    # scanner/mod.rs:51-66: `(x * 65536.0) as i32` -- truncation toward zero,
    # saturating at the i32 range, NaN -> 0 (Rust's float-to-int cast).
    f = O.lib().rro_fp_from_f64
    cases = [(0.3, 19660), (-0.3, -19660), (-1.5 / 65536, -1), (1.5 / 65536, 1), (-2.0, -131072),
             (-1.0 - 0.5 / 65536, -65536), (1e10, 2**31 - 1), (-1e10, -2**31), (float("nan"), 0)]
    for x, want in cases:
        assert f(x) == want, f"fp_from_f64({x}) = {f(x)}, want {want}"


# ---------------------------------------------------------------------------
# brute-force sampler
# ---------------------------------------------------------------------------

def brute_check(edges, eo, W, H):
    """Compares the oracle's AA coverage with the winding-number sampler on
    every cell outside the columns an edge touches in that AA row."""
    aa = [[Fraction(v) * 4 for v in e] for e in edges]
    s = O.Scanner(edges, eo, 0, 4 * H - 1, aa=True)
    aw = 4 * W
    for y in range(H):
        buf, _, _, _ = s.render_aa_line(aw, y)
        bits = np.unpackbits(buf, axis=1)[:, :aw]  # MSB-first, as AaBuf::set_span
        for k in range(4):
            r = 4 * y + k
            touched = np.zeros(aw, bool)
            cross = []
            for x0, y0, x1, y1 in aa:
                lo_y, hi_y = min(y0, y1), max(y0, y1)
                if not (int(np.floor(lo_y)) <= r <= int(np.floor(hi_y))):
                    continue
                if y0 == y1:
                    xa, xb = min(x0, x1), max(x0, x1)
                else:
                    ta, tb = max(lo_y, Fraction(r)), min(hi_y, Fraction(r + 1))
                    xs = [x0 + (t - y0) * (x1 - x0) / (y1 - y0) for t in (ta, tb)]
                    xa, xb = min(xs), max(xs)
                    if lo_y < r <= hi_y:
                        cross.append((x0 + (r - y0) * (x1 - x0) / (y1 - y0), 1 if y0 > y1 else -1))
                a = max(int(np.floor(xa)) - 1, 0)
                b = min(int(np.floor(xb)) + 1, aw - 1)
                if a <= b:
                    touched[a:b + 1] = True
            for c in range(aw):
                if touched[c]:
                    continue
                wind = sum(d for xr, d in cross if xr < c + Fraction(1, 2))
                inside = (wind % 2 != 0) if eo else (wind != 0)
                if bool(bits[k, c]) != inside:
                    raise AssertionError(f"AA cell ({c}, {r}): covered {bool(bits[k, c])}, sampler {inside}")


def check_brute_sampler(n=25, seed=5):
    rng = np.random.default_rng(seed)
    W, H = 9, 8
    for _ in range(n):
        e = random_polygon(rng, int(rng.integers(3, 6)), W, H)
        brute_check(e, int(rng.integers(0, 2)), W, H)


# ---------------------------------------------------------------------------
# the checks against the oracle
# ---------------------------------------------------------------------------

CHECKS = {
    "rect_documented": lambda: check_rectangles(O.MODE_DOCUMENTED),
    "rect_literal": lambda: check_rectangles(O.MODE_LITERAL),
    "rect_hairline": lambda: check_rectangles(O.MODE_DOCUMENTED, hairline=True),
    "row_counts_sum_to_zero": check_row_counts_sum_to_zero,
    "rows_sorted": check_rows_sorted,
    "nz_equals_eo_on_convex": check_nz_equals_eo_on_convex,
    "fp_truncates": check_fp_truncates,
    "brute_sampler": check_brute_sampler,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_passes(name):
    CHECKS[name]()


# mutant -> the check that must reject it (see RRO_MUTANT in oracle/rr_oracle.c)
MUTANTS = {
    1: "rect_documented",         # span end x1 instead of x1 + 1
    2: "rect_documented",         # nibble order swapped
    3: "rect_literal",            # literal mode draws the wrong AA row of the four
    4: "fp_truncates",            # floor instead of truncation
    5: "row_counts_sum_to_zero",  # a segment's first row keeps its count
    6: "rect_hairline",           # touching entries no longer merge
}


@pytest.fixture(scope="module")
def mutant_builds():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run(["make", "-s", "-C", os.path.join(root, "oracle"), "mutants"], check=True)


@pytest.mark.parametrize("m", sorted(MUTANTS))
def test_planted_mistake_is_rejected(mutant_builds, m):
    with O.using(O.mutant_path(m)):
        with pytest.raises(AssertionError):
            CHECKS[MUTANTS[m]]()


def test_brute_sampler_rejects_first_row_counts(mutant_builds):
    with O.using(O.mutant_path(5)):
        with pytest.raises(AssertionError):
            check_brute_sampler()
