// rr_host.cpp -- host side of the HIP backend: path ingestion (XPath::new +
// flatten_curve, done on the host in f64 exactly as the reference's GPU route
// does in gpu_fill_segs, crates/pdf_interp/src/renderer/page/gpu_ops.rs:34-100)
// and the seeded synthetic workloads the bench and tests run.
//
// Compiled with -ffp-contract=off: each f64 operation rounds as in the Rust
// source, and std::fma is used exactly where the Rust code calls mul_add.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "rr_hip.h"

namespace {

constexpr uint8_t kFirst = 1, kLast = 2, kClosed = 4, kCurve = 8;  // path/mod.rs:70-79
constexpr int kMaxCurveSplits = 1024;                                // types.rs MAX_CURVE_SPLITS

// f64::midpoint for page-range operands: (a + b) / 2.
inline double midpoint(double a, double b) {
    const double hi = 8.98846567431157953865e307;
    const double lo = 4.45014771701440227211e-308;
    const double aa = std::fabs(a), ab = std::fabs(b);
    if (aa <= hi && ab <= hi) return (a + b) / 2.0;
    if (aa < lo) return a + b / 2.0;
    if (ab < lo) return a / 2.0 + b;
    return a / 2.0 + b / 2.0;
}

// Adaptive De Casteljau flattening on the linked slot stack of
// crates/raster/src/path/flatten.rs:109-215.  Appends endpoints (p0 implicit).
struct CurveStack {
    double cx[(kMaxCurveSplits + 1) * 3];
    double cy[(kMaxCurveSplits + 1) * 3];
    int32_t next[kMaxCurveSplits + 1];
};

void flatten(const double p[8], double flat_sq, CurveStack &s, std::vector<double> &out) {
    s.cx[0] = p[0];
    s.cy[0] = p[1];
    s.cx[1] = p[2];
    s.cy[1] = p[3];
    s.cx[2] = p[4];
    s.cy[2] = p[5];
    s.cx[kMaxCurveSplits * 3] = p[6];
    s.cy[kMaxCurveSplits * 3] = p[7];
    s.next[0] = kMaxCurveSplits;
    int32_t a = 0;
    while (a < kMaxCurveSplits) {
        const int32_t b = s.next[a];
        const double x0 = s.cx[a * 3], y0 = s.cy[a * 3];
        const double x1 = s.cx[a * 3 + 1], y1 = s.cy[a * 3 + 1];
        const double x2 = s.cx[a * 3 + 2], y2 = s.cy[a * 3 + 2];
        const double x3 = s.cx[b * 3], y3 = s.cy[b * 3];
        const double mx = midpoint(x0, x3), my = midpoint(y0, y3);
        const double ax = x1 - mx, ay = y1 - my, bx = x2 - mx, by = y2 - my;
        const double d1 = std::fma(ax, ax, ay * ay);
        const double d2 = std::fma(bx, bx, by * by);
        if (b - a == 1 || (d1 <= flat_sq && d2 <= flat_sq)) {
            out.push_back(x3);
            out.push_back(y3);
            a = b;
            continue;
        }
        const double lx1 = midpoint(x0, x1), ly1 = midpoint(y0, y1);
        const double hx = midpoint(x1, x2), hy = midpoint(y1, y2);
        const double lx2 = midpoint(lx1, hx), ly2 = midpoint(ly1, hy);
        const double rx2 = midpoint(x2, x3), ry2 = midpoint(y2, y3);
        const double rx1 = midpoint(hx, rx2), ry1 = midpoint(hy, ry2);
        const double rx0 = midpoint(lx2, rx1), ry0 = midpoint(ly2, ry1);
        const int32_t c = (a + b) / 2;
        s.cx[a * 3 + 1] = lx1;
        s.cy[a * 3 + 1] = ly1;
        s.cx[a * 3 + 2] = lx2;
        s.cy[a * 3 + 2] = ly2;
        s.next[a] = c;
        s.cx[c * 3] = rx0;
        s.cy[c * 3] = ry0;
        s.cx[c * 3 + 1] = rx1;
        s.cy[c * 3 + 1] = ry1;
        s.cx[c * 3 + 2] = rx2;
        s.cy[c * 3 + 2] = ry2;
        s.next[c] = b;
    }
}

// XPath::new walk (crates/raster/src/xpath.rs:162-260) with close_subpaths =
// true as fill_impl passes it (fill/mod.rs:125).  A curve never checks the
// LAST flag, so a subpath ending in a curve chains into the next one.
void xpath_edges(const double *xy, const uint8_t *fl, int64_t n, const double m[6],
                 double flatness, std::vector<double> &edges) {
    const double flat_sq = flatness * flatness;
    std::vector<double> t(static_cast<size_t>(2 * n));
    for (int64_t i = 0; i < n; ++i) {  // transform (xpath.rs:421-426)
        const double xi = xy[2 * i], yi = xy[2 * i + 1];
        t[2 * i] = std::fma(xi, m[0], std::fma(yi, m[2], m[4]));
        t[2 * i + 1] = std::fma(xi, m[1], std::fma(yi, m[3], m[5]));
    }
    std::unique_ptr<CurveStack> stack;
    std::vector<double> flat;
    auto emit = [&](double x0, double y0, double x1, double y1) {
        edges.push_back(x0);
        edges.push_back(y0);
        edges.push_back(x1);
        edges.push_back(y1);
    };
    int64_t i = 0;
    while (i < n) {
        if (!(fl[i] & kFirst)) {
            ++i;
            continue;
        }
        const double sx = t[2 * i], sy = t[2 * i + 1];
        double cx = sx, cy = sy;
        ++i;
        while (i < n) {
            if (fl[i] & kCurve) {
                if (i + 2 >= n) break;
                const double p[8] = {cx, cy, t[2 * i], t[2 * i + 1], t[2 * i + 2], t[2 * i + 3],
                                     t[2 * i + 4], t[2 * i + 5]};
                if (!stack) stack.reset(new CurveStack);
                flat.clear();
                flatten(p, flat_sq, *stack, flat);
                for (size_t k = 0; k < flat.size(); k += 2) {
                    emit(cx, cy, flat[k], flat[k + 1]);
                    cx = flat[k];
                    cy = flat[k + 1];
                }
                i += 3;
            } else {
                const double nx = t[2 * i], ny = t[2 * i + 1];
                emit(cx, cy, nx, ny);
                cx = nx;
                cy = ny;
                const bool last = (fl[i] & kLast) != 0;
                ++i;
                if (last) break;
            }
        }
        if (std::fabs(cx - sx) > 1e-10 || std::fabs(cy - sy) > 1e-10) emit(cx, cy, sx, sy);
    }
}

// ---------------------------------------------------------------------------
// PathBuilder state machine (crates/raster/src/path/mod.rs:310-448)
// ---------------------------------------------------------------------------
struct Builder {
    std::vector<double> xy;
    std::vector<uint8_t> fl;
    size_t cur = 0;
    size_t n() const { return fl.size(); }
    void move_to(double x, double y) {
        const size_t len = n();
        xy.push_back(x);
        xy.push_back(y);
        fl.push_back(kFirst | kLast);
        cur = len;
    }
    void line_to(double x, double y) {
        fl.back() &= static_cast<uint8_t>(~kLast);
        xy.push_back(x);
        xy.push_back(y);
        fl.push_back(kLast);
    }
    void curve_to(double x1, double y1, double x2, double y2, double x3, double y3) {
        fl.back() &= static_cast<uint8_t>(~kLast);
        const double v[6] = {x1, y1, x2, y2, x3, y3};
        xy.insert(xy.end(), v, v + 6);
        fl.push_back(kCurve);
        fl.push_back(kCurve);
        fl.push_back(kLast);
    }
    void close(bool force) {
        const size_t sp = cur, last = n() - 1;
        const bool differ = !(xy[2 * sp] == xy[2 * last] && xy[2 * sp + 1] == xy[2 * last + 1]);
        if (force || (sp != last && differ)) line_to(xy[2 * sp], xy[2 * sp + 1]);
        fl[sp] |= kClosed;
        fl[n() - 1] |= kClosed;
        cur = n();
    }
};

// conic_to_cubic (crates/font/src/outline.rs:106-112)
void conic_to_cubic(double p0x, double p0y, double pcx, double pcy, double p3x, double p3y,
                    double out[4]) {
    out[0] = std::fma(p0x, 1.0 / 3.0, pcx * (2.0 / 3.0));
    out[1] = std::fma(p0y, 1.0 / 3.0, pcy * (2.0 / 3.0));
    out[2] = std::fma(pcx, 2.0 / 3.0, p3x * (1.0 / 3.0));
    out[3] = std::fma(pcy, 2.0 / 3.0, p3y * (1.0 / 3.0));
}

// ---------------------------------------------------------------------------
// Seeded generators
// ---------------------------------------------------------------------------
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * unit(); }
    int64_t range(int64_t lo, int64_t hi) {  // inclusive
        return lo + static_cast<int64_t>(next() % static_cast<uint64_t>(hi - lo + 1));
    }
    double log_uniform(double lo, double hi) { return std::exp(uniform(std::log(lo), std::log(hi))); }
};

struct Page {
    int32_t w = 0, h = 0;
    std::vector<double> pts;        // xy pairs of all paths
    std::vector<uint8_t> pt_flags;  // PathFlags
    std::vector<int64_t> pt_off{0};
    std::vector<double> edges;  // 4 per edge
    std::vector<int64_t> edge_off{0};
    std::vector<rr_path_attr> attrs;
    std::vector<int32_t> host_rows;

    void add_path(const Builder &b, uint8_t eo, uint8_t color, uint8_t a_input, double flatness) {
        pts.insert(pts.end(), b.xy.begin(), b.xy.end());
        pt_flags.insert(pt_flags.end(), b.fl.begin(), b.fl.end());
        pt_off.push_back(static_cast<int64_t>(pt_flags.size()));
        const size_t e0 = edges.size();
        static const double ident[6] = {1, 0, 0, 1, 0, 0};
        xpath_edges(b.xy.data(), b.fl.data(), static_cast<int64_t>(b.n()), ident, flatness, edges);
        edge_off.push_back(static_cast<int64_t>(edges.size() / 4));
        double ymn = INFINITY, ymx = -INFINITY;
        for (size_t k = e0; k < edges.size(); k += 4) {
            ymn = std::fmin(ymn, std::fmin(edges[k + 1], edges[k + 3]));
            ymx = std::fmax(ymx, std::fmax(edges[k + 1], edges[k + 3]));
        }
        int32_t r0 = 0, r1 = -1;
        if (ymn <= ymx) {
            r0 = static_cast<int32_t>(std::fmax(-1e9, std::floor(ymn)));
            r1 = static_cast<int32_t>(std::fmin(1e9, std::floor(ymx)));
        }
        host_rows.push_back(r0);
        host_rows.push_back(r1);
        attrs.push_back(rr_path_attr{eo, color, a_input, RR_ROUTE_SPAN});
    }
};

// Config 1: 500 random polygons, 3-64 vertices uniform over the page, NZ for
// even index and EO for odd, opaque grey U{0..255}, close(false).
void gen_polygons(Page &pg, SplitMix64 &rng, int n_paths) {
    for (int i = 0; i < n_paths; ++i) {
        const int k = static_cast<int>(rng.range(3, 64));
        Builder b;
        for (int v = 0; v < k; ++v) {
            const double x = rng.unit() * pg.w, y = rng.unit() * pg.h;
            if (v == 0)
                b.move_to(x, y);
            else
                b.line_to(x, y);
        }
        b.close(false);
        const uint8_t color = static_cast<uint8_t>(rng.range(0, 255));
        pg.add_path(b, static_cast<uint8_t>(i & 1), color, 255, 0.1);
    }
}

// Config 2: 6000 synthetic glyphs on a 60x100 grid of a Letter page.  Each
// glyph: bbox sides U[20,40] around the jittered cell centre, 1-3 closed
// contours of 12-40 quadratic (via conic_to_cubic) or cubic segments whose
// last segment ends exactly on the contour start, so close(false) adds no
// line and XPath::new chains the contours (curve-chaining).  NZ, black, opaque.
void gen_glyphs(Page &pg, SplitMix64 &rng, int rows, int cols, bool close_with_line) {
    const double pitch_x = static_cast<double>(pg.w) / cols, pitch_y = static_cast<double>(pg.h) / rows;
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const double cx = (c + 0.5) * pitch_x + rng.uniform(-2.0, 2.0);
            const double cy = (r + 0.5) * pitch_y + rng.uniform(-2.0, 2.0);
            const double gw = rng.uniform(20.0, 40.0), gh = rng.uniform(20.0, 40.0);
            const double bx0 = cx - gw / 2, by0 = cy - gh / 2, bx1 = cx + gw / 2, by1 = cy + gh / 2;
            auto clampx = [&](double x) { return std::fmin(bx1, std::fmax(bx0, x)); };
            auto clampy = [&](double y) { return std::fmin(by1, std::fmax(by0, y)); };
            Builder b;
            const int n_contours = static_cast<int>(rng.range(1, 3));
            for (int k = 0; k < n_contours; ++k) {
                const int nseg = static_cast<int>(rng.range(12, 40));
                const double scale = 1.0 - 0.3 * k;
                const double ox = cx + rng.uniform(-0.1, 0.1) * gw, oy = cy + rng.uniform(-0.1, 0.1) * gh;
                std::vector<double> on(2 * nseg);
                for (int s = 0; s < nseg; ++s) {
                    const double th = 6.283185307179586 * (s + rng.uniform(-0.3, 0.3)) / nseg;
                    const double rad = scale * rng.uniform(0.55, 1.0);
                    on[2 * s] = clampx(ox + 0.5 * gw * rad * std::cos(th));
                    on[2 * s + 1] = clampy(oy + 0.5 * gh * rad * std::sin(th));
                }
                b.move_to(on[0], on[1]);
                for (int s = 0; s < nseg; ++s) {
                    const int t = (s + 1) % nseg;
                    const double x0 = on[2 * s], y0 = on[2 * s + 1];
                    const double x3 = on[2 * t], y3 = on[2 * t + 1];
                    const double jx = 0.25 * gw, jy = 0.25 * gh;
                    if (rng.unit() < 0.5) {
                        const double pcx = clampx(0.5 * (x0 + x3) + rng.uniform(-jx, jx));
                        const double pcy = clampy(0.5 * (y0 + y3) + rng.uniform(-jy, jy));
                        double q[4];
                        conic_to_cubic(x0, y0, pcx, pcy, x3, y3, q);
                        b.curve_to(q[0], q[1], q[2], q[3], x3, y3);
                    } else {
                        const double c1x = clampx(x0 + (x3 - x0) / 3 + rng.uniform(-jx, jx));
                        const double c1y = clampy(y0 + (y3 - y0) / 3 + rng.uniform(-jy, jy));
                        const double c2x = clampx(x0 + 2 * (x3 - x0) / 3 + rng.uniform(-jx, jx));
                        const double c2y = clampy(y0 + 2 * (y3 - y0) / 3 + rng.uniform(-jy, jy));
                        b.curve_to(c1x, c1y, c2x, c2y, x3, y3);
                    }
                }
                if (close_with_line) b.line_to(on[0] + 0.5, on[1]);
                b.close(false);
            }
            pg.add_path(b, 0, 0, 255, 0.1);
        }
}

// Config 3 path: bbox sides log-uniform 4-4096 px (clamped to the page),
// 3-2000 edges log-uniform, star (smooth lobed outline sampled at n vertices)
// or zig-zag polygon inside the bbox, NZ with p=0.8, grey U{0..255}.
void gen_mixed_path(Page &pg, SplitMix64 &rng) {
    const double w = std::fmin(rng.log_uniform(4.0, 4096.0), pg.w - 1.0);
    const double h = std::fmin(rng.log_uniform(4.0, 4096.0), pg.h - 1.0);
    const double x0 = rng.uniform(0.0, pg.w - w), y0 = rng.uniform(0.0, pg.h - h);
    const int n = static_cast<int>(std::lround(rng.log_uniform(3.0, 2000.0)));
    const bool star = rng.unit() < 0.5;
    const uint8_t eo = rng.unit() < 0.2 ? 1 : 0;
    const uint8_t color = static_cast<uint8_t>(rng.range(0, 255));
    Builder b;
    if (star) {
        const int lobes = static_cast<int>(rng.range(2, 12));
        const double amp = rng.uniform(0.1, 0.6), phase = rng.uniform(0.0, 6.283185307179586);
        const double cx = x0 + w / 2, cy = y0 + h / 2;
        for (int k = 0; k < n; ++k) {
            const double th = 6.283185307179586 * k / n;
            const double r = 1.0 - amp + amp * (0.5 + 0.5 * std::sin(lobes * th + phase));
            const double x = cx + 0.5 * w * r * std::cos(th), y = cy + 0.5 * h * r * std::sin(th);
            if (k == 0)
                b.move_to(x, y);
            else
                b.line_to(x, y);
        }
    } else {
        for (int k = 0; k < n; ++k) {
            const double y = y0 + h * k / (n > 1 ? n - 1 : 1);
            const double x = (k & 1) ? x0 + w - rng.uniform(0.0, 0.35) * w : x0 + rng.uniform(0.0, 0.35) * w;
            if (k == 0)
                b.move_to(x, y);
            else
                b.line_to(x, y);
        }
    }
    b.close(false);
    pg.add_path(b, eo, color, 255, 0.1);
}

// ---------------------------------------------------------------------------
// Image workload: 32 images, RGB or CMYK, 1024-4096 px per side, placed by a
// rotation 0-30 deg and scale 0.5-2 onto Letter@300 pages under a smooth
// image-space soft mask; grey through a 17^4 (CMYK) or 33^3 (RGB) u16 grid.
// ---------------------------------------------------------------------------
struct Image {
    int32_t w = 0, h = 0, channels = 0, blend = 0, page = 0;
    double ctm[6] = {0, 0, 0, 0, 0, 0};
    std::vector<uint8_t> pixels, smask;
};

inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

void gen_image(Image &im, SplitMix64 &rng, int idx, int32_t page_w, int32_t page_h) {
    im.w = static_cast<int32_t>(rng.range(1024, 4096));
    im.h = static_cast<int32_t>(rng.range(1024, 4096));
    im.channels = (idx % 2 == 0) ? 4 : 3;
    im.blend = (idx % 2 == 0) ? 0 : 1;
    // content: per-channel gradients plus a little hashed noise
    const size_t n = static_cast<size_t>(im.w) * im.h;
    im.pixels.resize(n * im.channels);
    int32_t gx[4], gy[4], base[4];
    for (int c = 0; c < im.channels; ++c) {
        gx[c] = static_cast<int32_t>(rng.range(-300, 300));
        gy[c] = static_cast<int32_t>(rng.range(-300, 300));
        base[c] = static_cast<int32_t>(rng.range(0, 255));
    }
    const uint32_t salt = static_cast<uint32_t>(rng.next());
    for (int32_t y = 0; y < im.h; ++y)
        for (int32_t x = 0; x < im.w; ++x) {
            const size_t o = (static_cast<size_t>(y) * im.w + x) * im.channels;
            const uint32_t hsh = mix32(salt ^ static_cast<uint32_t>(y * 65599 + x));
            for (int c = 0; c < im.channels; ++c) {
                const int32_t v = base[c] + ((x * gx[c] + y * gy[c]) >> 10) + static_cast<int32_t>((hsh >> (8 * c)) & 15);
                im.pixels[o + c] = static_cast<uint8_t>(v & 255);
            }
        }
    // soft mask: integer bilinear upsampling of a 64x64 lattice of U{0..255}
    uint8_t lat[64][64];
    for (auto &row : lat)
        for (auto &v : row) v = static_cast<uint8_t>(rng.range(0, 255));
    im.smask.resize(n);
    for (int32_t y = 0; y < im.h; ++y) {
        const int64_t fy = static_cast<int64_t>(y) * 63 * 256 / (im.h - 1);
        const int32_t j = static_cast<int32_t>(fy >> 8), u = static_cast<int32_t>(fy & 255), j1 = j < 63 ? j + 1 : 63;
        for (int32_t x = 0; x < im.w; ++x) {
            const int64_t fx = static_cast<int64_t>(x) * 63 * 256 / (im.w - 1);
            const int32_t i = static_cast<int32_t>(fx >> 8), t = static_cast<int32_t>(fx & 255), i1 = i < 63 ? i + 1 : 63;
            const int32_t top = lat[j][i] * (256 - t) + lat[j][i1] * t;
            const int32_t bot = lat[j1][i] * (256 - t) + lat[j1][i1] * t;
            im.smask[static_cast<size_t>(y) * im.w + x] = static_cast<uint8_t>((top * (256 - u) + bot * u + 32768) >> 16);
        }
    }
    // placement: rotation, scale, centre uniform over the page (device y down);
    // the CTM maps the unit square in PDF space (y up) as the reference expects.
    double th, a, b, c, d;
    const double s = rng.uniform(0.5, 2.0);
    do {
        th = rng.uniform(0.0, 30.0) * 3.141592653589793 / 180.0;
        a = im.w * s * std::cos(th);
        b = im.w * s * std::sin(th);
        c = -im.h * s * std::sin(th);
        d = im.h * s * std::cos(th);
    } while (std::fabs(b) * im.w < 0.5 && std::fabs(c) * im.h < 0.5);  // keep the general affine path
    const double cx = rng.uniform(0.0, page_w), cy = page_h - rng.uniform(0.0, page_h);
    im.ctm[0] = a;
    im.ctm[1] = b;
    im.ctm[2] = c;
    im.ctm[3] = d;
    im.ctm[4] = cx - 0.5 * a - 0.5 * c;
    im.ctm[5] = cy - 0.5 * b - 0.5 * d;
}

// "Monotone-ish" grids: per axis a seeded non-decreasing curve (normalised
// cumulative sum of U[0,1) increments) combined per node, plus +-1% noise.
std::vector<uint16_t> gen_grid(SplitMix64 &rng, int dims, int g) {
    std::vector<std::vector<double>> curve(static_cast<size_t>(dims), std::vector<double>(static_cast<size_t>(g)));
    for (auto &cv : curve) {
        double acc = 0.0;
        cv[0] = 0.0;
        for (int i = 1; i < g; ++i) cv[static_cast<size_t>(i)] = (acc += rng.unit());
        for (auto &v : cv) v /= acc;
    }
    size_t n = 1;
    for (int a = 0; a < dims; ++a) n *= static_cast<size_t>(g);
    std::vector<uint16_t> grid(n);
    for (size_t idx = 0; idx < n; ++idx) {
        size_t r = idx;
        int ia[4] = {0, 0, 0, 0};
        // index layout: 4-D (k, c, m, y) with y fastest; 3-D (r, g, b) with b fastest
        for (int a = dims - 1; a >= 0; --a) {
            ia[a] = static_cast<int>(r % static_cast<size_t>(g));
            r /= static_cast<size_t>(g);
        }
        double v;
        if (dims == 4) {  // ia = {k, c, m, y}: reflectance falls with every ink
            const double k = curve[3][static_cast<size_t>(ia[0])], cc = curve[0][static_cast<size_t>(ia[1])];
            const double m = curve[1][static_cast<size_t>(ia[2])], y = curve[2][static_cast<size_t>(ia[3])];
            v = (1.0 - 0.3 * cc) * (1.0 - 0.59 * m) * (1.0 - 0.11 * y) * (1.0 - k);
        } else {  // ia = {r, g, b}: BT.709 weights
            v = 0.2126 * curve[0][static_cast<size_t>(ia[0])] + 0.7152 * curve[1][static_cast<size_t>(ia[1])] +
                0.0722 * curve[2][static_cast<size_t>(ia[2])];
        }
        v += rng.uniform(-0.01, 0.01);
        v = std::fmin(1.0, std::fmax(0.0, v));
        grid[idx] = static_cast<uint16_t>(std::lround(v * 65535.0));
    }
    return grid;
}

}  // namespace

struct rr_image_set {
    int32_t page_w = 2550, page_h = 3300, n_pages = 0, grid4_n = 17, grid3_n = 33;
    std::vector<Image> images;
    std::vector<uint16_t> grid4, grid3;
};

struct rr_workload {
    std::vector<Page> pages;
};

extern "C" {

int32_t rr_xpath_edges(const double *xy, const uint8_t *flags, int64_t n_pts,
                       const double matrix[6], double flatness, double *out_edges, int64_t cap,
                       int64_t *n_out) {
    if ((n_pts > 0 && (!xy || !flags)) || !matrix || !n_out || (cap > 0 && !out_edges))
        return RR_E_INVARIANT;
    std::vector<double> e;
    xpath_edges(xy, flags, n_pts, matrix, flatness, e);
    const int64_t n = static_cast<int64_t>(e.size() / 4);
    *n_out = n;
    const int64_t c = n < cap ? n : cap;
    if (c > 0) std::memcpy(out_edges, e.data(), static_cast<size_t>(c) * 4 * sizeof(double));
    return RR_OK;
}

int32_t rr_workload_create(int32_t config, uint64_t seed, int32_t max_pages, rr_workload **out) {
    if (!out) return RR_E_INVARIANT;
    auto w = std::make_unique<rr_workload>();
    SplitMix64 rng(seed);
    switch (config) {
        case 1: {
            Page pg;
            pg.w = 2550;
            pg.h = 3300;
            gen_polygons(pg, rng, 500);
            w->pages.push_back(std::move(pg));
            break;
        }
        case 2:
        case 102: {  // 102: the close-with-line variant of the glyph page
            Page pg;
            pg.w = 2550;
            pg.h = 3300;
            gen_glyphs(pg, rng, 60, 100, config == 102);
            w->pages.push_back(std::move(pg));
            break;
        }
        case 5: {
            // A stream of glyph pages drawn from the glyph-page distribution
            // (max_pages of them, 32 when 0), one generator across the stream.
            const int32_t n = max_pages > 0 ? max_pages : 32;
            for (int32_t k = 0; k < n; ++k) {
                Page pg;
                pg.w = 2550;
                pg.h = 3300;
                gen_glyphs(pg, rng, 60, 100, false);
                w->pages.push_back(std::move(pg));
            }
            break;
        }
        case 3: {
            // 40 Letter@300, 12 A3@300, 8 A4@600, 4 thumbnails@150, in this order.
            const int32_t dims[4][3] = {{40, 2550, 3300}, {12, 3508, 4961}, {8, 4960, 7016}, {4, 1275, 1650}};
            int made = 0;
            for (const auto &d : dims)
                for (int k = 0; k < d[0]; ++k) {
                    if (max_pages > 0 && made >= max_pages) break;
                    Page pg;
                    pg.w = d[1];
                    pg.h = d[2];
                    const int np = static_cast<int>(std::lround(rng.log_uniform(200.0, 20000.0)));
                    for (int i = 0; i < np; ++i) gen_mixed_path(pg, rng);
                    w->pages.push_back(std::move(pg));
                    ++made;
                }
            break;
        }
        default:
            return RR_E_INVARIANT;
    }
    *out = w.release();
    return RR_OK;
}

void rr_workload_destroy(rr_workload *w) { delete w; }

int32_t rr_workload_n_pages(const rr_workload *w) { return w ? static_cast<int32_t>(w->pages.size()) : 0; }

int32_t rr_workload_page_info(const rr_workload *w, int32_t p, int64_t info[5]) {
    if (!w || p < 0 || p >= static_cast<int32_t>(w->pages.size()) || !info) return RR_E_INVARIANT;
    const Page &pg = w->pages[static_cast<size_t>(p)];
    info[0] = pg.w;
    info[1] = pg.h;
    info[2] = static_cast<int64_t>(pg.attrs.size());
    info[3] = static_cast<int64_t>(pg.edges.size() / 4);
    info[4] = static_cast<int64_t>(pg.pt_flags.size());
    return RR_OK;
}

int32_t rr_workload_page_data(const rr_workload *w, int32_t p, const double **edges,
                              const int64_t **edge_off, const rr_path_attr **attrs,
                              const int32_t **host_rows, const double **pts,
                              const uint8_t **pt_flags, const int64_t **pt_off) {
    if (!w || p < 0 || p >= static_cast<int32_t>(w->pages.size())) return RR_E_INVARIANT;
    const Page &pg = w->pages[static_cast<size_t>(p)];
    if (edges) *edges = pg.edges.data();
    if (edge_off) *edge_off = pg.edge_off.data();
    if (attrs) *attrs = pg.attrs.data();
    if (host_rows) *host_rows = pg.host_rows.data();
    if (pts) *pts = pg.pts.data();
    if (pt_flags) *pt_flags = pg.pt_flags.data();
    if (pt_off) *pt_off = pg.pt_off.data();
    return RR_OK;
}

int32_t rr_images_create(uint64_t seed, int32_t n_images, int32_t images_per_page, rr_image_set **out) {
    if (!out || n_images <= 0 || images_per_page <= 0) return RR_E_INVARIANT;
    auto set = std::make_unique<rr_image_set>();
    SplitMix64 rng(seed);
    set->grid4 = gen_grid(rng, 4, set->grid4_n);
    set->grid3 = gen_grid(rng, 3, set->grid3_n);
    set->images.resize(static_cast<size_t>(n_images));
    for (int32_t i = 0; i < n_images; ++i) {
        gen_image(set->images[static_cast<size_t>(i)], rng, i, set->page_w, set->page_h);
        set->images[static_cast<size_t>(i)].page = i / images_per_page;
    }
    set->n_pages = (n_images + images_per_page - 1) / images_per_page;
    *out = set.release();
    return RR_OK;
}

void rr_images_destroy(rr_image_set *s) { delete s; }

int32_t rr_images_info(const rr_image_set *s, int32_t info[6]) {
    if (!s || !info) return RR_E_INVARIANT;
    info[0] = static_cast<int32_t>(s->images.size());
    info[1] = s->n_pages;
    info[2] = s->page_w;
    info[3] = s->page_h;
    info[4] = s->grid4_n;
    info[5] = s->grid3_n;
    return RR_OK;
}

int32_t rr_images_get(const rr_image_set *s, int32_t i, rr_image_desc *desc) {
    if (!s || !desc || i < 0 || i >= static_cast<int32_t>(s->images.size())) return RR_E_INVARIANT;
    const Image &im = s->images[static_cast<size_t>(i)];
    desc->pixels = im.pixels.data();
    desc->smask = im.smask.data();
    desc->width = im.w;
    desc->height = im.h;
    desc->channels = im.channels;
    desc->blend = im.blend;
    for (int k = 0; k < 6; ++k) desc->ctm[k] = im.ctm[k];
    desc->page = im.page;
    desc->reserved = 0;
    return RR_OK;
}

int32_t rr_images_grids(const rr_image_set *s, const uint16_t **grid4, const uint16_t **grid3) {
    if (!s) return RR_E_INVARIANT;
    if (grid4) *grid4 = s->grid4.data();
    if (grid3) *grid3 = s->grid3.data();
    return RR_OK;
}

uint64_t rr_cov_hash_term(int64_t pixel_index, uint8_t shape) {
    uint64_t z = ((static_cast<uint64_t>(pixel_index) << 8) | shape) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // extern "C"
