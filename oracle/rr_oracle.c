/*
 * rr_oracle.c -- CPU oracle for the rasterrocket GPU raster hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see rr_oracle.h).  A plain-C restatement of the
 * reference CPU rasteriser.  Built with -ffp-contract=off so that every f64
 * operation rounds exactly like the Rust source; fma() is used exactly where
 * the Rust code calls mul_add.
 *
 * Parity status:
 *   div255, splash_floor/ceil/round, AA_GAMMA, set_span, nibble coverage,
 *   add_segment (dxdy, dxdy_fp, flags), scanner merge/count_crossings,
 *   composite formulas: pinned by the reference's own known-answer tests,
 *   re-run in tests/test_oracle_known_answers.py.
 *   flatten_curve, XPath::new: pinned by the reference tests plus closed
 *   forms (tests/test_oracle_known_answers.py).
 *   Scanner construction, render_aa_line, the fill_impl AA loop (both row
 *   modes), draw_aa_line: pinned by quarter-pixel rectangle closed forms,
 *   row-count and sortedness invariants, non-zero = even-odd on convex
 *   polygons and a brute-force winding sampler, each shown to reject a
 *   planted mistake (RRO_MUTANT below; tests/test_oracle_invariants.py).
 *   Row sort: the reference sorts each row with Rust's sort_unstable_by_key
 *   (scanner/mod.rs:230).  For rows of at most 32 entries that sort is
 *   stable (insertion sort up to 20; the small-sort for <=32 is made of
 *   stable merges); above 32 entries its tie order on x0 is not reproduced
 *   here (stable order is used).  rro_scanner_stats reports how many rows are
 *   exposed to that ("parity unpinned" for rows > 32 entries with x0 ties).
 */
#define _POSIX_C_SOURCE 199309L  /* clock_gettime (rro_fill_page_timed) */
#include <time.h>

#include "rr_oracle.h"

/* RRO_MUTANT (default 0) plants one known mistake for the test-suite's
 * mutation check (tests/test_oracle_invariants.py), which builds each mutant
 * with `make mutants` and shows that the invariant checks reject it:
 *   1 render_aa_line ends a span at x1 instead of x1 + 1
 *   2 aa_coverage_span takes the nibbles in the wrong order
 *   3 literal row mode draws at aa_y % 4 == 2 instead of 3
 *   4 fp_from_f64 floors instead of truncating toward zero
 *   5 fill_intersections gives a segment's first row its count
 *   6 write_intersect merges overlapping entries only, not touching ones
 *   7 compute_a_src drops the soft mask when a shape is present
 *   8 Multiply blends the source with itself instead of the destination */
#ifndef RRO_MUTANT
#define RRO_MUTANT 0
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* colour / rounding primitives                                              */
/* ------------------------------------------------------------------------- */

/* crates/color/src/convert.rs:50-57 */
uint8_t rro_div255(uint32_t x) {
    uint32_t s = (x + (x >> 8) + 0x80u) >> 8;
    return (uint8_t)(s < 255u ? s : 255u);
}

/* crates/color/src/convert.rs:221-246 (saturate_f64_to_i32) */
static int32_t sat_f64_i32(double x) {
    if (!isfinite(x)) return (x == INFINITY) ? INT32_MAX : INT32_MIN;
    /* Rust `x as i64` saturates; the values reaching here are integral. */
    if (x >= 9223372036854775807.0) return INT32_MAX;
    if (x <= -9223372036854775808.0) return INT32_MIN;
    int64_t v = (int64_t)x;
    if (v > INT32_MAX) return INT32_MAX;
    if (v < INT32_MIN) return INT32_MIN;
    return (int32_t)v;
}

int32_t rro_splash_floor(double x) { return sat_f64_i32(floor(x)); } /* convert.rs:262 */
int32_t rro_splash_ceil(double x) { return sat_f64_i32(ceil(x)); }   /* convert.rs:284 */
int32_t rro_splash_round(double x) { return rro_splash_floor(x + 0.5); } /* convert.rs:305 */

/* crates/raster/src/scanner/mod.rs:51-66 */
int32_t rro_fp_from_f64(double x) {
    double fp = x * 65536.0;
    if (fp >= 2147483647.0) return INT32_MAX;
    if (fp <= -2147483648.0) return INT32_MIN;
    if (isnan(fp)) return 0; /* Rust `NaN as i32` == 0 */
#if RRO_MUTANT == 4
    return (int32_t)floor(fp);
#else
    return (int32_t)fp;      /* truncation toward zero */
#endif
}

/* crates/raster/src/fill/mod.rs:46-48 */
const uint8_t RRO_AA_GAMMA[17] = {0,   4,   11,  21,  32,  45,  59,  74, 90,
                                  108, 126, 145, 166, 187, 209, 231, 255};

static int32_t sat_add_i32(int32_t a, int32_t b) {
    int64_t s = (int64_t)a + (int64_t)b;
    if (s > INT32_MAX) return INT32_MAX;
    if (s < INT32_MIN) return INT32_MIN;
    return (int32_t)s;
}

/* Rust f64::min / f64::max: NaN operands are ignored (same as C fmin/fmax). */
static double fmin2(double a, double b) { return fmin(a, b); }
static double fmax2(double a, double b) { return fmax(a, b); }

static uint64_t f64_bits(double v) {
    uint64_t u;
    memcpy(&u, &v, 8);
    return u;
}

/* ------------------------------------------------------------------------- */
/* path builder -- crates/raster/src/path/mod.rs:291-451                     */
/* ------------------------------------------------------------------------- */

struct rro_path {
    double *xy;
    uint8_t *flags;
    int64_t n, cap;
    int64_t cur_subpath;
};

rro_path *rro_path_new(void) {
    rro_path *p = (rro_path *)calloc(1, sizeof(*p));
    return p;
}

void rro_path_free(rro_path *p) {
    if (!p) return;
    free(p->xy);
    free(p->flags);
    free(p);
}

static void path_push(rro_path *p, double x, double y, uint8_t f) {
    if (p->n == p->cap) {
        int64_t nc = p->cap ? p->cap * 2 : 16;
        p->xy = (double *)realloc(p->xy, (size_t)nc * 2 * sizeof(double));
        p->flags = (uint8_t *)realloc(p->flags, (size_t)nc);
        p->cap = nc;
    }
    p->xy[2 * p->n] = x;
    p->xy[2 * p->n + 1] = y;
    p->flags[p->n] = f;
    p->n++;
}

static int no_current_point(const rro_path *p) { return p->cur_subpath == p->n; }
static int one_point_subpath(const rro_path *p) { return p->n > 0 && p->cur_subpath == p->n - 1; }

/* move_to: path/mod.rs:310-319 */
int rro_path_move_to(rro_path *p, double x, double y) {
    if (one_point_subpath(p)) return 2;
    int64_t len = p->n;
    path_push(p, x, y, RRO_PATH_FIRST | RRO_PATH_LAST);
    p->cur_subpath = len;
    return 0;
}

/* line_to: path/mod.rs:334-344 */
int rro_path_line_to(rro_path *p, double x, double y) {
    if (no_current_point(p)) return 1;
    p->flags[p->n - 1] &= (uint8_t)~RRO_PATH_LAST;
    path_push(p, x, y, RRO_PATH_LAST);
    return 0;
}

/* curve_to: path/mod.rs:362-385 */
int rro_path_curve_to(rro_path *p, double x1, double y1, double x2, double y2, double x3,
                      double y3) {
    if (no_current_point(p)) return 1;
    p->flags[p->n - 1] &= (uint8_t)~RRO_PATH_LAST;
    path_push(p, x1, y1, RRO_PATH_CURVE);
    path_push(p, x2, y2, RRO_PATH_CURVE);
    path_push(p, x3, y3, RRO_PATH_LAST);
    return 0;
}

/* close: path/mod.rs:412-448.  `first != last` is PathPoint's derived
 * PartialEq, i.e. f64 == on both coordinates. */
int rro_path_close(rro_path *p, int force) {
    if (no_current_point(p)) return 1;
    int64_t sp = p->cur_subpath;
    int64_t last_idx = p->n - 1;
    double fx = p->xy[2 * sp], fy = p->xy[2 * sp + 1];
    double lx = p->xy[2 * last_idx], ly = p->xy[2 * last_idx + 1];
    int differ = !(fx == lx && fy == ly);
    if (force || (sp != last_idx && differ)) {
        int r = rro_path_line_to(p, fx, fy);
        if (r) return r;
    }
    int64_t new_last = p->n - 1;
    p->flags[sp] |= RRO_PATH_CLOSED;
    p->flags[new_last] |= RRO_PATH_CLOSED;
    p->cur_subpath = p->n;
    return 0;
}

int64_t rro_path_len(const rro_path *p) { return p->n; }

void rro_path_get(const rro_path *p, double *xy, uint8_t *flags) {
    if (p->n == 0) return;
    memcpy(xy, p->xy, (size_t)p->n * 2 * sizeof(double));
    memcpy(flags, p->flags, (size_t)p->n);
}

/* crates/font/src/outline.rs:106-112 */
void rro_conic_to_cubic(const double p0[2], const double pc[2], const double p3[2],
                        double p1_out[2], double p2_out[2]) {
    p1_out[0] = fma(p0[0], 1.0 / 3.0, pc[0] * (2.0 / 3.0));
    p1_out[1] = fma(p0[1], 1.0 / 3.0, pc[1] * (2.0 / 3.0));
    p2_out[0] = fma(pc[0], 2.0 / 3.0, p3[0] * (1.0 / 3.0));
    p2_out[1] = fma(pc[1], 2.0 / 3.0, p3[1] * (1.0 / 3.0));
}

/* ------------------------------------------------------------------------- */
/* flatten_curve -- crates/raster/src/path/flatten.rs:109-215                */
/* ------------------------------------------------------------------------- */

#define RRO_MAX_CURVE_SPLITS 1024 /* types.rs: MAX_CURVE_SPLITS */

/* f64::midpoint for operands within +-f64::MAX/2 is (a+b)/2. */
static double mid(double a, double b) {
    const double HI = 8.98846567431157953865e307; /* f64::MAX / 2 */
    const double LO = 4.45014771701440227211e-308; /* f64::MIN_POSITIVE * 2 */
    double aa = fabs(a), ab = fabs(b);
    if (aa <= HI && ab <= HI) return (a + b) / 2.0;
    if (aa < LO) return a + (b / 2.0);
    if (ab < LO) return (a / 2.0) + b;
    return (a / 2.0) + (b / 2.0);
}

typedef struct {
    double cx[(RRO_MAX_CURVE_SPLITS + 1) * 3];
    double cy[(RRO_MAX_CURVE_SPLITS + 1) * 3];
    int32_t cnext[RRO_MAX_CURVE_SPLITS + 1];
} curve_data;

/* Appends flattened endpoints (p0 implicit) via callback-free buffer. */
typedef struct {
    double *xy;
    int64_t n, cap;
} ptbuf;

static void ptbuf_push(ptbuf *b, double x, double y) {
    if (b->n == b->cap) {
        int64_t nc = b->cap ? b->cap * 2 : 64;
        b->xy = (double *)realloc(b->xy, (size_t)nc * 2 * sizeof(double));
        b->cap = nc;
    }
    b->xy[2 * b->n] = x;
    b->xy[2 * b->n + 1] = y;
    b->n++;
}

static void flatten_impl(const double p[8], double flatness_sq, ptbuf *out, curve_data *d) {
    const int32_t max = RRO_MAX_CURVE_SPLITS;
    d->cx[0] = p[0];
    d->cy[0] = p[1];
    d->cx[1] = p[2];
    d->cy[1] = p[3];
    d->cx[2] = p[4];
    d->cy[2] = p[5];
    d->cx[max * 3] = p[6];
    d->cy[max * 3] = p[7];
    d->cnext[0] = max;
    int32_t pp1 = 0;
    while (pp1 < max) {
        int32_t pp2 = d->cnext[pp1];
        double xl0 = d->cx[pp1 * 3], yl0 = d->cy[pp1 * 3];
        double xx1 = d->cx[pp1 * 3 + 1], yy1 = d->cy[pp1 * 3 + 1];
        double xx2 = d->cx[pp1 * 3 + 2], yy2 = d->cy[pp1 * 3 + 2];
        double xr3 = d->cx[pp2 * 3], yr3 = d->cy[pp2 * 3];
        double mx = mid(xl0, xr3), my = mid(yl0, yr3);
        double dx1 = xx1 - mx, dy1 = yy1 - my, dx2 = xx2 - mx, dy2 = yy2 - my;
        double d1 = fma(dx1, dx1, dy1 * dy1);
        double d2 = fma(dx2, dx2, dy2 * dy2);
        if (pp2 - pp1 == 1 || (d1 <= flatness_sq && d2 <= flatness_sq)) {
            ptbuf_push(out, xr3, yr3);
            pp1 = pp2;
        } else {
            double xl1 = mid(xl0, xx1), yl1 = mid(yl0, yy1);
            double xh = mid(xx1, xx2), yh = mid(yy1, yy2);
            double xl2 = mid(xl1, xh), yl2 = mid(yl1, yh);
            double xr2 = mid(xx2, xr3), yr2 = mid(yy2, yr3);
            double xr1 = mid(xh, xr2), yr1 = mid(yh, yr2);
            double xr0 = mid(xl2, xr1), yr0 = mid(yl2, yr1);
            int32_t pp3 = (pp1 + pp2) / 2; /* i32::midpoint, both >= 0 */
            d->cx[pp1 * 3 + 1] = xl1;
            d->cy[pp1 * 3 + 1] = yl1;
            d->cx[pp1 * 3 + 2] = xl2;
            d->cy[pp1 * 3 + 2] = yl2;
            d->cnext[pp1] = pp3;
            d->cx[pp3 * 3] = xr0;
            d->cy[pp3 * 3] = yr0;
            d->cx[pp3 * 3 + 1] = xr1;
            d->cy[pp3 * 3 + 1] = yr1;
            d->cx[pp3 * 3 + 2] = xr2;
            d->cy[pp3 * 3 + 2] = yr2;
            d->cnext[pp3] = pp2;
        }
    }
}

int64_t rro_flatten_curve(const double p[8], double flatness_sq, double *out_xy, int64_t cap) {
    curve_data *d = (curve_data *)malloc(sizeof(curve_data));
    ptbuf b = {0, 0, 0};
    flatten_impl(p, flatness_sq, &b, d);
    int64_t n = b.n;
    for (int64_t i = 0; i < n && i < cap; ++i) {
        out_xy[2 * i] = b.xy[2 * i];
        out_xy[2 * i + 1] = b.xy[2 * i + 1];
    }
    free(b.xy);
    free(d);
    return n;
}

/* ------------------------------------------------------------------------- */
/* XPath -- crates/raster/src/xpath.rs                                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    double *e; /* 4 per edge */
    int64_t n, cap;
} edgebuf;

static void edge_push(edgebuf *b, double x0, double y0, double x1, double y1) {
    if (b->n == b->cap) {
        int64_t nc = b->cap ? b->cap * 2 : 64;
        b->e = (double *)realloc(b->e, (size_t)nc * 4 * sizeof(double));
        b->cap = nc;
    }
    double *q = b->e + 4 * b->n;
    q[0] = x0;
    q[1] = y0;
    q[2] = x1;
    q[3] = y1;
    b->n++;
}

/* XPath::new (xpath.rs:162-260) with transform (xpath.rs:421-426).  Stroke
 * adjustment hints are absent on fill paths (build_adjusts gets no hints). */
static void xpath_walk(const double *xy, const uint8_t *flags, int64_t n, const double m[6],
                       double flatness, int close_subpaths, edgebuf *out) {
    double flatness_sq = flatness * flatness;
    double *t = (double *)malloc((size_t)(n > 0 ? n : 1) * 2 * sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        double xi = xy[2 * i], yi = xy[2 * i + 1];
        t[2 * i] = fma(xi, m[0], fma(yi, m[2], m[4]));
        t[2 * i + 1] = fma(xi, m[1], fma(yi, m[3], m[5]));
    }
    curve_data *cd = NULL;
    ptbuf fb = {0, 0, 0};
    int64_t i = 0;
    while (i < n) {
        if (flags[i] & RRO_PATH_FIRST) {
            double sp_x = t[2 * i], sp_y = t[2 * i + 1];
            double cur_x = sp_x, cur_y = sp_y;
            i += 1;
            while (i < n) {
                if (flags[i] & RRO_PATH_CURVE) {
                    if (i + 2 >= n) break;
                    double p[8] = {cur_x,        cur_y,        t[2 * i],       t[2 * i + 1],
                                   t[2 * i + 2], t[2 * i + 3], t[2 * i + 4],   t[2 * i + 5]};
                    if (!cd) cd = (curve_data *)malloc(sizeof(curve_data));
                    fb.n = 0;
                    flatten_impl(p, flatness_sq, &fb, cd);
                    for (int64_t k = 0; k < fb.n; ++k) {
                        edge_push(out, cur_x, cur_y, fb.xy[2 * k], fb.xy[2 * k + 1]);
                        cur_x = fb.xy[2 * k];
                        cur_y = fb.xy[2 * k + 1];
                    }
                    i += 3;
                } else {
                    double nx = t[2 * i], ny = t[2 * i + 1];
                    edge_push(out, cur_x, cur_y, nx, ny);
                    cur_x = nx;
                    cur_y = ny;
                    int is_last = (flags[i] & RRO_PATH_LAST) != 0;
                    i += 1;
                    if (is_last) break;
                }
            }
            if (close_subpaths && (fabs(cur_x - sp_x) > 1e-10 || fabs(cur_y - sp_y) > 1e-10))
                edge_push(out, cur_x, cur_y, sp_x, sp_y);
        } else {
            i += 1;
        }
    }
    free(fb.xy);
    free(cd);
    free(t);
}

int64_t rro_xpath_edges(const double *xy, const uint8_t *flags, int64_t n_pts, const double m[6],
                        double flatness, int close_subpaths, double *out_edges, int64_t cap) {
    edgebuf b = {0, 0, 0};
    xpath_walk(xy, flags, n_pts, m, flatness, close_subpaths, &b);
    int64_t n = b.n;
    int64_t c = n < cap ? n : cap;
    if (c > 0) memcpy(out_edges, b.e, (size_t)c * 4 * sizeof(double));
    free(b.e);
    return n;
}

/* XPath::add_segment (xpath.rs:324-402) */
static rro_seg add_segment(double x0, double y0, double x1, double y1) {
    rro_seg s;
    uint32_t flags = 0;
    if (f64_bits(y0) == f64_bits(y1)) {
        flags |= RRO_SEG_HORIZ;
        if (f64_bits(x0) == f64_bits(x1)) flags |= RRO_SEG_VERT;
        s.x0 = x0;
        s.y0 = y0;
        s.x1 = x1;
        s.y1 = y1;
        s.dxdy = 0.0;
        s.dxdy_fp = 0;
        s.flags = flags;
        return s;
    }
    if (f64_bits(x0) == f64_bits(x1)) flags |= RRO_SEG_VERT;
    double dxdy = (flags & RRO_SEG_VERT) ? 0.0 : (x1 - x0) / (y1 - y0);
    if (y0 > y1) {
        double tx = x0, ty = y0;
        x0 = x1;
        y0 = y1;
        x1 = tx;
        y1 = ty;
        flags |= RRO_SEG_FLIPPED;
    }
    int32_t fp;
    double f = dxdy * 65536.0;
    if (f >= 2147483647.0)
        fp = INT32_MAX;
    else if (f <= -2147483648.0)
        fp = INT32_MIN;
    else if (isnan(f))
        fp = 0;
    else
        fp = (int32_t)round(f); /* f64::round: half away from zero */
    s.x0 = x0;
    s.y0 = y0;
    s.x1 = x1;
    s.y1 = y1;
    s.dxdy = dxdy;
    s.dxdy_fp = fp;
    s.flags = flags;
    return s;
}

void rro_add_segments(const double *edges, int64_t n, rro_seg *out) {
    for (int64_t i = 0; i < n; ++i)
        out[i] = add_segment(edges[4 * i], edges[4 * i + 1], edges[4 * i + 2], edges[4 * i + 3]);
}

/* XPath::aa_scale (xpath.rs:283-303): coordinates x4, dxdy unchanged. */
void rro_aa_scale(rro_seg *segs, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        segs[i].x0 *= 4.0;
        segs[i].y0 *= 4.0;
        segs[i].x1 *= 4.0;
        segs[i].y1 *= 4.0;
    }
}

/* ------------------------------------------------------------------------- */
/* XPathScanner -- crates/raster/src/scanner/mod.rs                          */
/* ------------------------------------------------------------------------- */

struct rro_scanner {
    int eo;
    int32_t x_min, x_max, y_min, y_max;
    int64_t n_rows;
    uint32_t *row_start; /* n_rows + 1 */
    rro_intersect *ints;
    int64_t rows_gt32, rows_gt32_tie;
};

/* write_intersect (scanner/mod.rs:600-642) */
static void write_intersect(int32_t row, int32_t y_min, const uint32_t *alloc_start,
                            uint32_t *cursors, rro_intersect *ints, int32_t x0, int32_t x1,
                            int32_t count) {
    int32_t lo = x0 < x1 ? x0 : x1, hi = x0 < x1 ? x1 : x0;
    int64_t idx = (int64_t)row - y_min;
    uint32_t base = alloc_start[idx];
    uint32_t cur = cursors[idx];
    if (cur > 0) {
        rro_intersect *last = &ints[base + cur - 1];
#if RRO_MUTANT == 6
        if (last->x1 >= lo && last->x0 <= hi) {
#else
        if (sat_add_i32(last->x1, 1) >= lo && last->x0 <= sat_add_i32(hi, 1)) {
#endif
            last->count = (int32_t)((uint32_t)last->count + (uint32_t)count);
            if (lo < last->x0) last->x0 = lo;
            if (hi > last->x1) last->x1 = hi;
            return;
        }
    }
    ints[base + cur].x0 = lo;
    ints[base + cur].x1 = hi;
    ints[base + cur].count = count;
    cursors[idx] = cur + 1;
}

/* Stable sort by x0: insertion sort for short rows, merge sort above. */
static void sort_row(rro_intersect *a, int64_t n, rro_intersect *tmp) {
    if (n <= 32) {
        for (int64_t i = 1; i < n; ++i) {
            rro_intersect v = a[i];
            int64_t j = i;
            while (j > 0 && a[j - 1].x0 > v.x0) {
                a[j] = a[j - 1];
                --j;
            }
            a[j] = v;
        }
        return;
    }
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t m = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = m, k = lo;
            while (i < m && j < hi) tmp[k++] = (a[j].x0 < a[i].x0) ? a[j++] : a[i++];
            while (i < m) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, (size_t)n * sizeof(rro_intersect));
    }
}

/* Row range of a segment (scanner/mod.rs:481-484, 523-526). */
static void seg_rows(const rro_seg *s, int32_t y_min, int32_t y_max, int32_t *r0, int32_t *r1) {
    double sy_min = fmin2(s->y0, s->y1), sy_max = fmax2(s->y0, s->y1);
    int32_t a = rro_splash_floor(sy_min), b = rro_splash_floor(sy_max);
    *r0 = a > y_min ? a : y_min;
    *r1 = b < y_max ? b : y_max;
}

rro_scanner *rro_scanner_new(const rro_seg *segs, int64_t n, int eo, int32_t clip_y_min,
                             int32_t clip_y_max) {
    rro_scanner *s = (rro_scanner *)calloc(1, sizeof(*s));
    s->eo = eo;
    s->x_min = 1;
    s->x_max = 0;
    s->y_min = 0;
    s->y_max = -1;
    if (n == 0 || clip_y_min > clip_y_max) return s;
    /* bbox over non-NaN segments overlapping the clip range (:138-160) */
    double xmn = INFINITY, xmx = -INFINITY, ymn = INFINITY, ymx = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const rro_seg *g = &segs[i];
        if (isnan(g->x0) || isnan(g->y0) || isnan(g->x1) || isnan(g->y1)) continue;
        double sy_min = fmin2(g->y0, g->y1), sy_max = fmax2(g->y0, g->y1);
        if (sy_min >= (double)clip_y_max + 1.0 || sy_max < (double)clip_y_min) continue;
        xmn = fmin2(fmin2(xmn, g->x0), g->x1);
        xmx = fmax2(fmax2(xmx, g->x0), g->x1);
        ymn = fmin2(ymn, sy_min);
        ymx = fmax2(ymx, sy_max);
    }
    if (xmn > xmx) return s;
    int32_t x_min = rro_splash_floor(xmn), x_max = rro_splash_floor(xmx);
    int32_t y_min = rro_splash_floor(ymn), y_max = rro_splash_floor(ymx);
    if (y_min < clip_y_min) y_min = clip_y_min;
    if (y_max > clip_y_max) y_max = clip_y_max;
    if (y_min > y_max) return s;
    int64_t n_rows = (int64_t)y_max - y_min + 1;
    uint32_t *counts = (uint32_t *)calloc((size_t)n_rows, sizeof(uint32_t));
    /* pass 1: count_intersections (:476-499) */
    for (int64_t i = 0; i < n; ++i) {
        const rro_seg *g = &segs[i];
        if (isnan(g->x0) || isnan(g->y0)) continue;
        int32_t r0, r1;
        seg_rows(g, y_min, y_max, &r0, &r1);
        if (g->flags & RRO_SEG_HORIZ) {
            int32_t row = rro_splash_floor(fmin2(g->y0, g->y1));
            if (row >= y_min && row <= y_max) counts[row - y_min]++;
        } else {
            for (int64_t r = r0; r <= r1; ++r) counts[r - y_min]++;
        }
    }
    /* pass 2: prefix sum */
    uint32_t *alloc_start = (uint32_t *)malloc((size_t)(n_rows + 1) * sizeof(uint32_t));
    uint64_t total = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        alloc_start[r] = (uint32_t)total;
        total += counts[r];
    }
    alloc_start[n_rows] = (uint32_t)total;
    rro_intersect *ints = (rro_intersect *)calloc((size_t)(total ? total : 1), sizeof(rro_intersect));
    memset(counts, 0, (size_t)n_rows * sizeof(uint32_t));
    /* pass 3: fill_intersections (:511-587) */
    for (int64_t i = 0; i < n; ++i) {
        const rro_seg *g = &segs[i];
        if (isnan(g->x0) || isnan(g->y0)) continue;
        double seg_y_min = fmin2(g->y0, g->y1);
        int32_t r0, r1;
        seg_rows(g, y_min, y_max, &r0, &r1);
        if (g->flags & RRO_SEG_HORIZ) {
            int32_t row = rro_splash_floor(seg_y_min);
            if (row >= y_min && row <= y_max) {
                int32_t x0 = rro_splash_floor(fmin2(g->x0, g->x1));
                int32_t x1 = rro_splash_floor(fmax2(g->x0, g->x1));
                write_intersect(row, y_min, alloc_start, counts, ints, x0, x1, 0);
            }
        } else if (g->flags & RRO_SEG_VERT) {
            int32_t cnt = (g->flags & RRO_SEG_FLIPPED) ? 1 : -1;
            int32_t x = rro_splash_floor(g->x0);
            for (int64_t r = r0; r <= r1; ++r) {
#if RRO_MUTANT == 5
                int32_t c = cnt;
#else
                int32_t c = (seg_y_min < (double)r) ? cnt : 0;
#endif
                write_intersect((int32_t)r, y_min, alloc_start, counts, ints, x, x, c);
            }
        } else {
            int32_t cnt = (g->flags & RRO_SEG_FLIPPED) ? 1 : -1;
            double x_base = fma(g->y0, -g->dxdy, g->x0);
            double sx_min = fmin2(g->x0, g->x1), sx_max = fmax2(g->x0, g->x1);
            int32_t fpmin = rro_fp_from_f64(sx_min), fpmax = rro_fp_from_f64(sx_max);
            double xx_init = ((double)r0 > g->y0) ? fma((double)r0, g->dxdy, x_base) : g->x0;
            double cl = xx_init;
            if (cl < sx_min) cl = sx_min;
            if (cl > sx_max) cl = sx_max;
            int32_t px0 = rro_splash_floor(cl);
            int32_t xx1_fp = rro_fp_from_f64(fma((double)r0 + 1.0, g->dxdy, x_base));
            for (int64_t r = r0; r <= r1; ++r) {
                int32_t c1 = xx1_fp < fpmin ? fpmin : (xx1_fp > fpmax ? fpmax : xx1_fp);
                int32_t px1 = c1 >> 16;
#if RRO_MUTANT == 5
                int32_t c = cnt;
#else
                int32_t c = (seg_y_min < (double)r) ? cnt : 0;
#endif
                write_intersect((int32_t)r, y_min, alloc_start, counts, ints, px0, px1, c);
                px0 = px1;
                xx1_fp = sat_add_i32(xx1_fp, g->dxdy_fp);
            }
        }
    }
    /* pass 4: sort each row by x0, compact (:221-237) */
    uint32_t *row_start = (uint32_t *)malloc((size_t)(n_rows + 1) * sizeof(uint32_t));
    rro_intersect *tmp = (rro_intersect *)malloc((size_t)(total ? total : 1) * sizeof(rro_intersect));
    uint64_t wp = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        row_start[r] = (uint32_t)wp;
        uint32_t base = alloc_start[r], nw = counts[r];
        if (nw > 32) {
            s->rows_gt32++;
            int tie = 0;
            for (uint32_t a = 0; a < nw && !tie; ++a)
                for (uint32_t b = a + 1; b < nw; ++b)
                    if (ints[base + a].x0 == ints[base + b].x0) {
                        tie = 1;
                        break;
                    }
            s->rows_gt32_tie += tie;
        }
        sort_row(ints + base, nw, tmp);
        memmove(ints + wp, ints + base, (size_t)nw * sizeof(rro_intersect));
        wp += nw;
    }
    row_start[n_rows] = (uint32_t)wp;
    free(tmp);
    free(alloc_start);
    free(counts);
    s->x_min = x_min;
    s->x_max = x_max;
    s->y_min = y_min;
    s->y_max = y_max;
    s->n_rows = n_rows;
    s->row_start = row_start;
    s->ints = ints;
    return s;
}

void rro_scanner_free(rro_scanner *s) {
    if (!s) return;
    free(s->row_start);
    free(s->ints);
    free(s);
}

static int scanner_empty(const rro_scanner *s) { return s->x_min > s->x_max || s->y_min > s->y_max; }

void rro_scanner_bbox(const rro_scanner *s, int32_t out[4]) {
    out[0] = s->x_min;
    out[1] = s->x_max;
    out[2] = s->y_min;
    out[3] = s->y_max;
}

/* XPathScanner::row (:295-318) */
static const rro_intersect *scanner_row(const rro_scanner *s, int32_t y, int64_t *n) {
    if (scanner_empty(s) || y < s->y_min || y > s->y_max) {
        *n = 0;
        return NULL;
    }
    int64_t i = (int64_t)y - s->y_min;
    *n = (int64_t)s->row_start[i + 1] - s->row_start[i];
    return s->ints + s->row_start[i];
}

int64_t rro_scanner_row(const rro_scanner *s, int32_t y, rro_intersect *out, int64_t cap) {
    int64_t n;
    const rro_intersect *r = scanner_row(s, y, &n);
    for (int64_t i = 0; i < n && i < cap; ++i) out[i] = r[i];
    return n;
}

/* XPathScanner::test (:330-352) */
int rro_scanner_test(const rro_scanner *s, int32_t x, int32_t y) {
    int64_t n;
    const rro_intersect *r = scanner_row(s, y, &n);
    int32_t m = s->eo ? 1 : -1;
    int32_t count = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (r[i].x0 > x) break;
        if (x <= r[i].x1) return 1;
        count = (int32_t)((uint32_t)count + (uint32_t)r[i].count);
    }
    return (count & m) != 0;
}

void rro_scanner_stats(const rro_scanner *s, int64_t out[4]) {
    out[0] = s->n_rows;
    out[1] = s->n_rows ? (int64_t)s->row_start[s->n_rows] : 0;
    out[2] = s->rows_gt32;
    out[3] = s->rows_gt32_tie;
}

/* AaBuf::set_span (bitmap.rs:535-574), on one row of ceil(width/8) bytes. */
void rro_aabuf_set_span(uint8_t *row, int32_t width, int32_t x0, int32_t x1) {
    if (x0 > width) x0 = width;
    if (x1 > width) x1 = width;
    if (x0 >= x1) return;
    int32_t b0 = x0 >> 3, b1 = (x1 - 1) >> 3;
    uint8_t rmask = (x1 & 7) == 0 ? 0xff : (uint8_t)~(0xffu >> (x1 & 7));
    if (b0 == b1) {
        row[b0] |= (uint8_t)((0xffu >> (x0 & 7)) & rmask);
    } else {
        row[b0] |= (uint8_t)(0xffu >> (x0 & 7));
        for (int32_t b = b0 + 1; b < b1; ++b) row[b] = 0xff;
        row[b1] |= rmask;
    }
}

/* render_aa_line (scanner/mod.rs:393-467).  The reference loops forever when
 * a row's winding never returns to zero; here that row stops after the span
 * to aa_width (counted in the return value). */
int rro_scanner_render_aa_line(const rro_scanner *s, uint8_t *aabuf, int32_t aa_width,
                               int32_t *x0, int32_t *x1, int32_t y) {
    int32_t rb = (aa_width + 7) / 8;
    memset(aabuf, 0, (size_t)rb * 4);
    int32_t xx_min = aa_width, xx_max = -1;
    int32_t m = s->eo ? 1 : -1;
    int unbalanced = 0;
    for (int32_t yy = 0; yy < 4; ++yy) {
        int64_t scan_y = (int64_t)y * 4 + yy;
        int64_t n = 0;
        const rro_intersect *row =
            (scan_y >= INT32_MIN && scan_y <= INT32_MAX) ? scanner_row(s, (int32_t)scan_y, &n) : NULL;
        int32_t count = 0;
        int64_t i = 0;
        int32_t xx = 0;
        while (i < n || (count & m) != 0) {
            if ((count & m) == 0) {
                if (i >= n) break;
                xx = row[i].x0;
                count = (int32_t)((uint32_t)count + (uint32_t)row[i].count);
                i++;
                continue;
            }
            int32_t span_x1;
            int hit_end = 0;
            for (;;) {
                if (i >= n) {
                    span_x1 = aa_width;
                    hit_end = 1;
                    break;
                }
                count = (int32_t)((uint32_t)count + (uint32_t)row[i].count);
#if RRO_MUTANT == 1
                int32_t end = row[i].x1;
#else
                int32_t end = sat_add_i32(row[i].x1, 1);
#endif
                i++;
                if ((count & m) == 0) {
                    span_x1 = end;
                    break;
                }
            }
            int32_t bx0 = xx > 0 ? xx : 0;
            int32_t bx1 = span_x1 < aa_width ? span_x1 : aa_width;
            if (bx1 < 0) bx1 = 0;
            if (bx0 < bx1) {
                rro_aabuf_set_span(aabuf + (size_t)yy * rb, aa_width, bx0, bx1);
                if (bx0 < xx_min) xx_min = bx0;
                if (bx1 > xx_max) xx_max = bx1;
            }
            xx = span_x1;
            if (hit_end) {
                unbalanced++;
                break;
            }
        }
    }
    if (xx_min > xx_max) {
        *x0 = 0;
        *x1 = -1;
    } else {
        *x0 = xx_min / 4;
        *x1 = (xx_max - 1) / 4;
    }
    return unbalanced;
}

/* aa_coverage_span_scalar (simd/aa_coverage.rs:38,51-69) */
void rro_aa_coverage_span(const uint8_t *aabuf, int32_t row_bytes, int32_t x0, int32_t count,
                          uint8_t *shape) {
    static const uint8_t NIBBLE_POP[16] = {0, 1, 1, 2, 1, 2, 2, 3, 1, 2, 2, 3, 2, 3, 3, 4};
    for (int32_t i = 0; i < count; ++i) {
        int32_t x = x0 + i;
        int32_t bi = x >> 1;
        uint8_t c = 0;
        for (int r = 0; r < 4; ++r) {
            uint8_t b = aabuf[(size_t)r * row_bytes + bi];
#if RRO_MUTANT == 2
            c += NIBBLE_POP[(x & 1) ? (b >> 4) : (b & 0x0f)];
#else
            c += NIBBLE_POP[(x & 1) ? (b & 0x0f) : (b >> 4)];
#endif
        }
        shape[i] = c;
    }
}

/* ------------------------------------------------------------------------- */
/* compositing                                                               */
/* ------------------------------------------------------------------------- */

/* render_span_aa_inner, no-alpha arm (pipe/aa.rs:192-208), one channel. */
uint8_t rro_composite_gray8(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape) {
    uint32_t a = rro_div255((uint32_t)a_input * shape);
    return rro_div255((255u - a) * dst + a * src);
}

/* composite_aa_rgb8_opaque (simd/composite.rs:30-32,56-114), one channel. */
uint8_t rro_composite_rgb8_fast(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape) {
    uint32_t a = ((uint32_t)a_input * shape + 255u) >> 8;
    return (uint8_t)(((255u - a) * dst + a * src + 255u) >> 8);
}

/* rgb_to_gray (crates/pdf_raster/src/render.rs:1306-1324) */
int32_t rro_rgb_to_gray(int32_t r, int32_t g, int32_t b) {
    return (2126 * r + 7152 * g + 722 * b + 5000) / 10000;
}

uint64_t rro_cov_hash_term(int64_t pixel_index, uint8_t shape) {
    uint64_t z = (((uint64_t)pixel_index) << 8 | shape) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* ------------------------------------------------------------------------- */
/* fill -- crates/raster/src/fill/mod.rs:110-244, draw_aa_line :344-401      */
/* ------------------------------------------------------------------------- */

typedef struct {
    double x_min, y_min, x_max, y_max;
    int32_t x_min_i, y_min_i, x_max_i, y_max_i;
} clip_rect;

/* Clip::new + recompute_int_bounds (clip.rs:73-88,266-271) */
static clip_rect make_clip(const double c[4]) {
    clip_rect r;
    r.x_min = fmin2(c[0], c[2]);
    r.x_max = fmax2(c[0], c[2]);
    r.y_min = fmin2(c[1], c[3]);
    r.y_max = fmax2(c[1], c[3]);
    r.x_min_i = rro_splash_floor(r.x_min);
    r.y_min_i = rro_splash_floor(r.y_min);
    r.x_max_i = rro_splash_ceil(r.x_max) - 1;
    r.y_max_i = rro_splash_ceil(r.y_max) - 1;
    return r;
}

enum { CLIP_ALL_INSIDE, CLIP_ALL_OUTSIDE, CLIP_PARTIAL };

/* Clip::test_rect (clip.rs:196-220), no path clips. */
static int clip_test_rect(const clip_rect *c, int32_t l, int32_t t, int32_t r, int32_t b) {
    if ((double)r + 1.0 <= c->x_min || (double)l >= c->x_max || (double)b + 1.0 <= c->y_min ||
        (double)t >= c->y_max)
        return CLIP_ALL_OUTSIDE;
    if ((double)l >= c->x_min && (double)r + 1.0 <= c->x_max && (double)t >= c->y_min &&
        (double)b + 1.0 <= c->y_max)
        return CLIP_ALL_INSIDE;
    return CLIP_PARTIAL;
}

/* compute_a_src (pipe/general.rs:326-344) for a fill (shape present). */
uint32_t rro_compute_a_src(uint32_t a_input, uint32_t soft, uint32_t shape, int has_soft_mask, int has_shape) {
#if RRO_MUTANT == 7
    if (has_soft_mask && !has_shape) return rro_div255(a_input * soft);
#else
    if (has_soft_mask)
        return has_shape ? rro_div255((uint32_t)rro_div255(a_input * soft) * shape) : rro_div255(a_input * soft);
#endif
    return has_shape ? rro_div255(a_input * shape) : a_input;
}

/* render_span_general_inner, no-alpha arm (pipe/general.rs:270-315), one
 * channel: Normal c = div255((255 - a_src) d + a_src s); any other mode
 * div255((255 - a_src) d + a_src c_b) with Multiply's c_b = div255(s d)
 * (pipe/blend.rs:24-26). */
uint8_t rro_composite_general(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape, uint8_t soft,
                              int has_soft_mask, int multiply) {
    uint32_t a_src = rro_compute_a_src(a_input, soft, shape, has_soft_mask, 1);
#if RRO_MUTANT == 8
    uint32_t cb = multiply ? rro_div255((uint32_t)src * src) : src;
#else
    uint32_t cb = multiply ? rro_blend_multiply(src, dst) : src;
#endif
    return rro_div255((255u - a_src) * dst + a_src * cb);
}

/* The paint of one fill: render_span's choice (pipe/mod.rs:157-199) between
 * the AA path (no soft mask, Normal) and the general pipe, for a 1-byte
 * (grey, or Rgb8 with r = g = b) or a 3-byte Rgb8 canvas. */
typedef struct {
    uint32_t flags;      /* RRO_FILL_* */
    uint8_t color[3];    /* r, g, b (a grey fill repeats its value) */
    uint8_t a_input;
    int flavour;
    const uint8_t *soft; /* page-space soft mask (RRO_FILL_SOFT_MASK) */
    int32_t soft_stride;
} rro_paint;

static uint8_t paint_channel(const rro_paint *pt, uint8_t d, uint8_t src, uint8_t sh, uint8_t sm) {
    if (pt->flags & (RRO_FILL_MULTIPLY | RRO_FILL_SOFT_MASK))
        return rro_composite_general(d, src, pt->a_input, sh, sm, (pt->flags & RRO_FILL_SOFT_MASK) != 0,
                                     (pt->flags & RRO_FILL_MULTIPLY) != 0);
    if (pt->flavour == RRO_FLAVOUR_GRAY8) return rro_composite_gray8(d, src, pt->a_input, sh);
    return rro_composite_rgb8_fast(d, src, pt->a_input, sh);
}

/* draw_aa_line (fill/mod.rs:344-401): shapes of [x0, x1], skip when all are
 * zero, else render_span over the whole range. */
static int64_t draw_aa_line(uint8_t *canvas, int32_t w, int32_t stride, const uint8_t *aabuf,
                            int32_t rb, int32_t x0, int32_t x1, int32_t y, const rro_paint *pt,
                            uint8_t *shape_buf, uint8_t *shape_out, uint64_t *cov_hash) {
    int32_t count = x1 - x0 + 1;
    rro_aa_coverage_span(aabuf, rb, x0, count, shape_buf);
    int any = 0;
    for (int32_t i = 0; i < count; ++i)
        if (shape_buf[i]) {
            shape_buf[i] = RRO_AA_GAMMA[shape_buf[i]];
            any = 1;
        }
    if (!any) return 0;
    int64_t drawn = 0;
    uint8_t *row = canvas + (size_t)y * stride;
    const int nch = pt->flavour == RRO_FLAVOUR_RGB8 ? 3 : 1;
    uint64_t hsum = 0;
    for (int32_t i = 0; i < count; ++i) {
        uint8_t sh = shape_buf[i];
        int32_t x = x0 + i;
        uint8_t sm = (pt->flags & RRO_FILL_SOFT_MASK) ? pt->soft[(size_t)y * pt->soft_stride + x] : 255;
        for (int ch = 0; ch < nch; ++ch)
            row[nch * x + ch] = paint_channel(pt, row[nch * x + ch], pt->color[ch], sh, sm);
        if (sh) {
            drawn++;
            if (shape_out) shape_out[(size_t)y * w + x] = sh;
            if (cov_hash) hsum += rro_cov_hash_term((int64_t)y * w + x, sh);
        }
    }
    if (cov_hash) *cov_hash += hsum;
    return drawn;
}

static int64_t fill_segs(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, rro_seg *segs,
                         int64_t n, int eo, const rro_paint *pt, int mode,
                         const clip_rect *clip, uint8_t *shape_out, uint64_t *cov_hash) {
    if (n == 0) return 0;
    rro_aa_scale(segs, n);
    int32_t cy0 = clip->y_min_i * 4;
    int32_t cy1 = (clip->y_max_i + 1) * 4 - 1;
    rro_scanner *s = rro_scanner_new(segs, n, eo, cy0, cy1);
    int64_t drawn = 0;
    if (scanner_empty(s) || w == 0) {
        rro_scanner_free(s);
        return 0;
    }
    int cres = clip_test_rect(clip, s->x_min / 4, s->y_min / 4, s->x_max / 4, s->y_max / 4);
    if (cres == CLIP_ALL_OUTSIDE) {
        rro_scanner_free(s);
        return 0;
    }
    int32_t aw = w * 4;
    int32_t rb = (aw + 7) / 8;
    uint8_t *aabuf = (uint8_t *)calloc((size_t)rb * 4, 1);
    uint8_t *shape_buf = (uint8_t *)malloc((size_t)w);
    if (mode == RRO_MODE_LITERAL) {
        for (int64_t aa_y = s->y_min; aa_y <= s->y_max; ++aa_y) {
            int32_t y = (int32_t)(aa_y / 4);
            int32_t aa_row = (int32_t)(aa_y % 4);
            int32_t x0 = s->x_min / 4, x1 = s->x_max / 4;
            rro_scanner_render_aa_line(s, aabuf, aw, &x0, &x1, (int32_t)aa_y);
            if (cres != CLIP_ALL_INSIDE) {
                if (x0 < clip->x_min_i) x0 = clip->x_min_i;
                if (x1 > clip->x_max_i) x1 = clip->x_max_i;
            }
#if RRO_MUTANT == 3
            if (aa_row == 2) {
#else
            if (aa_row == 3) {
#endif
                if (x0 <= x1 && y >= 0 && y < h)
                    drawn += draw_aa_line(canvas, w, stride, aabuf, rb, x0, x1, y, pt, shape_buf, shape_out,
                                          cov_hash);
                memset(aabuf, 0, (size_t)rb * 4);
            }
        }
    } else {
        for (int32_t y = s->y_min / 4; y <= s->y_max / 4; ++y) {
            int32_t x0 = s->x_min / 4, x1 = s->x_max / 4;
            rro_scanner_render_aa_line(s, aabuf, aw, &x0, &x1, y);
            if (cres != CLIP_ALL_INSIDE) {
                if (x0 < clip->x_min_i) x0 = clip->x_min_i;
                if (x1 > clip->x_max_i) x1 = clip->x_max_i;
            }
            if (x0 <= x1 && y >= 0 && y < h)
                drawn += draw_aa_line(canvas, w, stride, aabuf, rb, x0, x1, y, pt, shape_buf, shape_out, cov_hash);
        }
    }
    free(shape_buf);
    free(aabuf);
    rro_scanner_free(s);
    return drawn;
}

int64_t rro_fill_ex(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges, int64_t n_edges,
                    uint32_t flags, const uint8_t color[3], uint8_t a_input, int mode, int flavour,
                    const double clip[4], const uint8_t *soft_mask, int32_t soft_stride, uint8_t *shape_out,
                    uint64_t *cov_hash) {
    if (n_edges <= 0) return 0;
    rro_paint pt = {flags, {color[0], color[1], color[2]}, a_input, flavour, soft_mask, soft_stride};
    rro_seg *segs = (rro_seg *)malloc((size_t)n_edges * sizeof(rro_seg));
    rro_add_segments(edges, n_edges, segs);
    clip_rect c = make_clip(clip);
    int64_t d = fill_segs(canvas, w, h, stride, segs, n_edges, (flags & RRO_FILL_EVEN_ODD) != 0, &pt, mode, &c,
                          shape_out, cov_hash);
    free(segs);
    return d;
}

int64_t rro_fill(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                 int64_t n_edges, int eo, uint8_t color, uint8_t a_input, int mode, int flavour,
                 const double clip[4], uint8_t *shape_out, uint64_t *cov_hash) {
    const uint8_t c3[3] = {color, color, color};
    return rro_fill_ex(canvas, w, h, stride, edges, n_edges, eo ? RRO_FILL_EVEN_ODD : 0u, c3, a_input, mode, flavour,
                       clip, NULL, 0, shape_out, cov_hash);
}

int64_t rro_fill_page_ex(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                         const int64_t *edge_off, int64_t n_paths, const uint8_t *flags, const uint8_t *colors,
                         const uint8_t *a_input, int mode, int flavour, const uint8_t *soft_mask,
                         int32_t soft_stride, uint64_t *cov_hashes) {
    double clip[4] = {0.0, 0.0, (double)w - 0.001, (double)h - 0.001};
    int64_t total = 0;
    for (int64_t p = 0; p < n_paths; ++p) {
        int64_t off = edge_off[p], n = edge_off[p + 1] - edge_off[p];
        total += rro_fill_ex(canvas, w, h, stride, edges + 4 * off, n, flags[p], colors + 3 * p, a_input[p], mode,
                             flavour, clip, soft_mask, soft_stride, NULL, cov_hashes ? &cov_hashes[p] : NULL);
    }
    return total;
}

int64_t rro_fill_page(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                      const int64_t *edge_off, int64_t n_paths, const uint8_t *eo,
                      const uint8_t *color, const uint8_t *a_input, int mode, int flavour,
                      uint64_t *cov_hashes) {
    double clip[4] = {0.0, 0.0, (double)w - 0.001, (double)h - 0.001};
    int64_t total = 0;
    for (int64_t p = 0; p < n_paths; ++p) {
        int64_t off = edge_off[p], n = edge_off[p + 1] - edge_off[p];
        total += rro_fill(canvas, w, h, stride, edges + 4 * off, n, eo[p], color[p], a_input[p],
                          mode, flavour, clip, NULL, cov_hashes ? &cov_hashes[p] : NULL);
    }
    return total;
}

/* rro_fill_page with each path's wall time (CLOCK_MONOTONIC, ns) in ns[p]:
 * the per-path CPU cost the dispatcher bench prices the CPU class with. */
int64_t rro_fill_page_timed(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                            const int64_t *edge_off, int64_t n_paths, const uint8_t *eo, const uint8_t *color,
                            const uint8_t *a_input, int mode, int64_t *ns) {
    double clip[4] = {0.0, 0.0, (double)w - 0.001, (double)h - 0.001};
    int64_t total = 0;
    for (int64_t p = 0; p < n_paths; ++p) {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        int64_t off = edge_off[p], n = edge_off[p + 1] - edge_off[p];
        total += rro_fill(canvas, w, h, stride, edges + 4 * off, n, eo[p], color[p], a_input[p], mode,
                          RRO_FLAVOUR_GRAY8, clip, NULL, NULL);
        clock_gettime(CLOCK_MONOTONIC, &t1);
        ns[p] = (int64_t)(t1.tv_sec - t0.tv_sec) * 1000000000 + (t1.tv_nsec - t0.tv_nsec);
    }
    return total;
}

/* rgb_to_gray over a whole Rgb8 bitmap (render.rs:1306-1324). */
void rro_rgb_to_gray_page(const uint8_t *rgb, int32_t rgb_stride, uint8_t *gray, int32_t gray_stride, int32_t w,
                          int32_t h) {
    for (int32_t y = 0; y < h; ++y)
        for (int32_t x = 0; x < w; ++x) {
            const uint8_t *p = rgb + (size_t)y * rgb_stride + 3 * (size_t)x;
            gray[(size_t)y * gray_stride + x] = (uint8_t)rro_rgb_to_gray(p[0], p[1], p[2]);
        }
}

/* The reference dispatcher's pixel bbox area (gpu_fill_segs,
 * crates/pdf_interp/src/renderer/page/gpu_ops.rs:52-84) over an edge list:
 * f64 min / max of the endpoints (NaN skipped, as Rust's f64::min / max),
 * clamped to [0, w] x [0, h], floored / ceiled; -1 when not finite or empty.
 * A fill goes to the GPU when this is at least 256 (crates/gpu/src/lib.rs:
 * 112,120; gpu_ops.rs:242,289). */
int64_t rro_ref_bbox_area(const double *edges, int64_t n, int32_t w, int32_t h) {
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int64_t i = 0; i < n; ++i) {
        const double *e = edges + 4 * i;
        x0 = fmin(fmin(x0, e[0]), e[2]);
        y0 = fmin(fmin(y0, e[1]), e[3]);
        x1 = fmax(fmax(x1, e[0]), e[2]);
        y1 = fmax(fmax(y1, e[1]), e[3]);
    }
    if (!isfinite(x0) || !isfinite(y0) || !isfinite(x1) || !isfinite(y1)) return -1;
    x0 = floor(fmax(x0, 0.0));
    y0 = floor(fmax(y0, 0.0));
    x1 = ceil(fmin(x1, (double)w));
    y1 = ceil(fmin(y1, (double)h));
    if (x1 <= x0 || y1 <= y0) return -1;
    return (int64_t)(x1 - x0) * (int64_t)(y1 - y0);
}
