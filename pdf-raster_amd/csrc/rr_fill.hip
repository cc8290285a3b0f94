// rr_fill.hip -- gfx950 kernels for the anti-aliased scanline fill +
// composite hot path, and the C-ABI backend that records them.
//
// Semantics are those of the reference CPU rasteriser, bit for bit:
//   edge setup      XPath::add_segment + aa_scale (crates/raster/src/xpath.rs:283-402)
//                   and the first-row setup of fill_intersections
//                   (crates/raster/src/scanner/mod.rs:511-587)
//   entries         per AA row, edges in path order, merged into the last
//                   entry (write_intersect, scanner/mod.rs:600-642)
//   row sort        stable by x0 (scanner/mod.rs:230)
//   winding walk    render_aa_line (scanner/mod.rs:393-467)
//   coverage        4 AA rows -> nibble popcount -> AA_GAMMA
//                   (fill/mod.rs:46-48,344-401; simd/aa_coverage.rs:51-69)
//   composite       Gray8 exact div255 (pipe/aa.rs:192-208) or the Rgb8
//                   opaque fast path (simd/composite.rs:56-114)
//
// Work decomposition: one wave64 owns a band of 16 output rows = 64 AA rows,
// one lane per AA row.  The wave walks the path's edge records in order 64 at
// a time (coalesced 32 B records staged in LDS), ballots which overlap the
// band, and every lane folds its own row's intersection into its entry list
// (LDS, [slot][lane] so that lanes never conflict).  Entries whose winding
// contribution cannot change the inside/outside state (count 0, or even under
// even-odd) are dropped when they are retired: they never start or end a span.
// The walk turns each row into disjoint sorted intervals; the wave then sweeps
// the band 256 pixels at a time through an LDS bitmap (one row per lane, no
// atomics) and composites 4 pixels per lane per dword, touching memory only
// for dwords that hold covered pixels.
//
// Launch shapes: PLAIN launches one grid per path (one wave per band), in
// path order on the context stream.  BATCHED launches one persistent grid for
// every page of the batch; each wave pulls (page, band) work items from an
// atomic queue and composites, in path order, every path of that page that
// overlaps its band.  Bands are disjoint, so the in-order compositing of
// overlapping paths holds by construction without any inter-wave ordering.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rr_hip.h"
#include "rr_internal.h"

#include "rr_device.h"

using namespace rrdev;

namespace {
// ---------------------------------------------------------------------------
// kernel 1: per-path edge setup + scanner bounding box
// ---------------------------------------------------------------------------
// One workgroup per path.  Each thread takes edges of the path, applies
// add_segment (flags, dxdy computed before the swap, dxdy_fp = round-half-away
// of dxdy*65536 saturated) and aa_scale (x4, exact), and emits the integer
// record of fill_intersections' first row.  The block reduces the scanner
// bbox over the non-NaN segments that overlap the AA clip rows.
// Vertical and horizontal edges whose columns fit 16 bits become sloped
// records with no slope (b = d = e = column << 16: every row gives the same
// column; a horizontal edge has one row, whose count is 0, so its entry is
// {a, b} as before), and a sloped record whose 16.16 stepping stays inside
// int32 over its rows is marked EDGE_FAST: row_entry then steps it in 32-bit
// arithmetic, which equals the saturating 64-bit form there.
__device__ __forceinline__ void to_fast_form(EdgeRec &r) {
    if (r.r0 > r.r1) return;
    const int kind = r.meta & 3;
    const bool fits_a = r.a >= -32768 && r.a <= 32767;
    if (kind == KIND_VERT && fits_a) {
        r.meta = (r.meta & ~3) | KIND_SLOPED;
        r.b = r.d = r.e = static_cast<int32_t>(static_cast<uint32_t>(r.a) << 16);
        r.c = 0;
    } else if (kind == KIND_HORIZ && fits_a && r.b >= -32768 && r.b <= 32767) {
        r.meta = KIND_SLOPED | 8;
        r.b = r.d = r.e = static_cast<int32_t>(static_cast<uint32_t>(r.b) << 16);
        r.c = 0;
    }
    if ((r.meta & 3) == KIND_SLOPED) {
        const long long last = static_cast<long long>(r.b) + static_cast<long long>(r.r1 - r.r0) * r.c;
        if (last >= INT32_MIN && last <= INT32_MAX) r.meta |= EDGE_FAST;
    }
}

// Area of the reference dispatcher's pixel bbox (gpu_ops.rs:52-84): clamped to
// the bitmap, floored / ceiled to whole pixels; -1 when the bbox is not finite
// or empty (the reference then keeps the fill on the CPU), saturated at
// INT32_MAX.
__device__ __forceinline__ int32_t ref_bbox_area(double x0, double y0, double x1, double y1, int32_t w, int32_t h) {
    if (!isfinite(x0) || !isfinite(y0) || !isfinite(x1) || !isfinite(y1)) return -1;
    x0 = floor(fmax(x0, 0.0));
    y0 = floor(fmax(y0, 0.0));
    x1 = ceil(fmin(x1, static_cast<double>(w)));
    y1 = ceil(fmin(y1, static_cast<double>(h)));
    if (x1 <= x0 || y1 <= y0) return -1;
    const double a = (x1 - x0) * (y1 - y0);
    return a >= 2147483647.0 ? INT32_MAX : static_cast<int32_t>(a);
}

// Persistent waves (four per block) over the paths: for each path,
// add_segment + aa_scale of every edge into its EdgeRec, the scanner bbox, the
// chunk summaries (min first row, max last row of every global 64-edge chunk:
// a plain store for a chunk inside the path, atomics for one it shares with a
// neighbour; csum starts at {INT32_MAX, INT32_MIN}), and the dispatcher's
// features.  Long-lived waves keep the chip full where one short wave per path
// would leave it waiting on wave launches, and the page of a path is found by
// a forward scan of the pages' first paths kept in LDS.
constexpr int kSetupPagesLds = 1024;  // pages whose first path index the block keeps in LDS
__device__ __forceinline__ void edge_setup_path(const double *__restrict__ edges,
                                                const int64_t *__restrict__ edge_off, const int64_t p,
                                                const DevPage *__restrict__ pages, const int32_t lo,
                                                EdgeRec *__restrict__ recs, PathHdr *__restrict__ hdrs,
                                                int2 *__restrict__ rows, int4 *__restrict__ sel,
                                                int32_t *__restrict__ nb, const rr_path_attr *__restrict__ attrs,
                                                float *__restrict__ feat, int2 *__restrict__ csum, int32_t mode,
                                                int32_t *__restrict__ err) {
    const int lane = threadIdx.x & 63;
    const DevPage &pg = pages[lo];
    const ClipRect clip = pg.clip;
    const int64_t e0 = edge_off[p], e1 = edge_off[p + 1];
    const int32_t cy0 = clip.y_min_i * 4, cy1 = (clip.y_max_i + 1) * 4 - 1;
    int32_t xmn = INT32_MAX, xmx = INT32_MIN, ymn = INT32_MAX, ymx = INT32_MIN;
    int any = 0;
    double shoelace = 0.0, dy_sum = 0.0;  // dispatcher features
    // the reference dispatcher's pixel bbox over the segment endpoints
    // (gpu_fill_segs, crates/pdf_interp/src/renderer/page/gpu_ops.rs:52-84):
    // f64 min / max, which skip NaN like Rust's f64::min / max
    double rx0 = INFINITY, ry0 = INFINITY, rx1 = -INFINITY, ry1 = -INFINITY;
    // the next chunk's edges are loaded while this one is converted
    auto load_edge = [&](int64_t i) {
        double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
        if (i >= e0 && i < e1) q = reinterpret_cast<const double4 *>(edges)[i];
        return q;
    };
    double4 qn = load_edge((e0 & ~63ll) + lane);
    for (int64_t base = e0 & ~63ll; base < e1; base += 64) {
        const int64_t i = base + lane;
        int32_t cr0 = INT32_MAX, cr1 = INT32_MIN;
        const double4 q = qn;
        if (base + 64 < e1) qn = load_edge(base + 64 + lane);
        if (i >= e0 && i < e1) {
            double x0 = q.x, y0 = q.y, x1 = q.z, y1 = q.w;
            rx0 = fmin(fmin(rx0, x0), x1);
            ry0 = fmin(fmin(ry0, y0), y1);
            rx1 = fmax(fmax(rx1, x0), x1);
            ry1 = fmax(fmax(ry1, y0), y1);
            if (feat && isfinite(x0) && isfinite(y0) && isfinite(x1) && isfinite(y1)) {
                shoelace += x0 * y1 - x1 * y0;
                dy_sum += fabs(y1 - y0);
            }
            const bool horiz = bits_of(y0) == bits_of(y1);
            const bool vert = bits_of(x0) == bits_of(x1);
            double dxdy = 0.0;
            int32_t dxdy_fp = 0;
            bool flipped = false;
            if (!horiz) {
                if (!vert) dxdy = (x1 - x0) / (y1 - y0);
                if (y0 > y1) {
                    double t = x0;
                    x0 = x1;
                    x1 = t;
                    t = y0;
                    y0 = y1;
                    y1 = t;
                    flipped = true;
                }
                dxdy_fp = cvt_i32_sat(round(dxdy * 65536.0));  // saturated, NaN -> 0
            }
            x0 *= 4.0;
            y0 *= 4.0;
            x1 *= 4.0;
            y1 *= 4.0;
            const double sy_min = fmin(y0, y1), sy_max = fmax(y0, y1);
            if (!(isnan(x0) || isnan(y0) || isnan(x1) || isnan(y1)) &&
                !(sy_min >= static_cast<double>(cy1) + 1.0 || sy_max < static_cast<double>(cy0))) {
                any = 1;
                xmn = min(xmn, splash_floor(fmin(x0, x1)));
                xmx = max(xmx, splash_floor(fmax(x0, x1)));
                ymn = min(ymn, splash_floor(sy_min));
                ymx = max(ymx, splash_floor(sy_max));
            }
            EdgeRec r{1, 0, 0, 0, 0, 0, 0, 0};
            if (!(isnan(x0) || isnan(y0))) {
                const int32_t fy0 = splash_floor(sy_min), fy1 = splash_floor(sy_max);
                if (horiz) {
                    if (fy0 >= cy0 && fy0 <= cy1) {
                        r.r0 = r.r1 = fy0;
                        r.meta = KIND_HORIZ;
                        r.a = splash_floor(fmin(x0, x1));
                        r.b = splash_floor(fmax(x0, x1));
                    }
                } else {
                    r.r0 = max(fy0, cy0);
                    r.r1 = min(fy1, cy1);
                    const int32_t first_zero = (sy_min < static_cast<double>(r.r0)) ? 0 : 8;
                    const int32_t pos = flipped ? 4 : 0;
                    if (vert) {
                        r.meta = KIND_VERT | pos | first_zero;
                        r.a = splash_floor(x0);
                    } else {
                        r.meta = KIND_SLOPED | pos | first_zero;
                        const double x_base = fma(y0, -dxdy, x0);
                        const double sx_min = fmin(x0, x1), sx_max = fmax(x0, x1);
                        const double rr0 = static_cast<double>(r.r0);
                        double xx = (rr0 > y0) ? fma(rr0, dxdy, x_base) : x0;
                        if (xx < sx_min) xx = sx_min;
                        if (xx > sx_max) xx = sx_max;
                        r.a = splash_floor(xx);
                        r.b = fp_from_f64(fma(rr0 + 1.0, dxdy, x_base));
                        r.c = dxdy_fp;
                        r.d = fp_from_f64(sx_min);
                        r.e = fp_from_f64(sx_max);
                    }
                }
                to_fast_form(r);
            }
            recs[i] = r;
            if (r.r0 <= r.r1) {
                cr0 = r.r0;
                cr1 = r.r1;
            }
        }
        cr0 = wave_min(cr0);
        cr1 = wave_max(cr1);
        if (lane == 0) {
            const int64_t ch = base >> 6;
            if (base >= e0 && base + 64 <= e1) {
                csum[ch] = make_int2(cr0, cr1);
            } else {
                atomicMin(&csum[ch].x, cr0);
                atomicMax(&csum[ch].y, cr1);
            }
        }
    }
    xmn = wave_min(xmn);
    xmx = wave_max(xmx);
    ymn = wave_min(ymn);
    ymx = wave_max(ymx);
    any = __ballot(any != 0) != 0;
    for (int o = 32; o > 0; o >>= 1) {
        rx0 = fmin(rx0, __shfl_xor(rx0, o));
        ry0 = fmin(ry0, __shfl_xor(ry0, o));
        rx1 = fmax(rx1, __shfl_xor(rx1, o));
        ry1 = fmax(ry1, __shfl_xor(ry1, o));
    }
    if (feat) {
        shoelace = wave_sum_f64(shoelace);
        dy_sum = wave_sum_f64(dy_sum);
    }
    if (lane == 0) {
        if ((attrs[p].eo & RR_FILL_SOFT_MASK) && !pg.smask) atomicOr(err, kErrNoSoftMask);
        PathHdr h{0, -1, 1, 0, HDR_EMPTY, lo, 0, 0};
        h.ref_area = ref_bbox_area(rx0, ry0, rx1, ry1, pg.width, pg.height);
        int2 drawn = make_int2(1, 0);
        if (any) {
            const int32_t y_min = max(ymn, cy0), y_max = min(ymx, cy1);
            if (y_min <= y_max) {
                h.y_min = y_min;
                h.y_max = y_max;
                h.x_min = xmn;
                h.x_max = xmx;
                h.flags = 0;
                // Clip::test_rect on the pixel bbox (clip.rs:196-220; fill/mod.rs:156-170)
                const int32_t l = h.x_min / 4, t = h.y_min / 4, rt = h.x_max / 4, b = h.y_max / 4;
                if (static_cast<double>(rt) + 1.0 <= clip.x_min || static_cast<double>(l) >= clip.x_max ||
                    static_cast<double>(b) + 1.0 <= clip.y_min || static_cast<double>(t) >= clip.y_max)
                    h.flags = HDR_OUTSIDE;
                else
                    drawn = drawn_rows(h, mode, pg.height);
            }
        }
        hdrs[p] = h;
        rows[p] = drawn;
        // pixel columns that can hold coverage (widened to the clip's right
        // edge later for a path with an unbalanced row), and its band count
        int4 s4 = make_int4(1, 0, 1, 0);
        int32_t bands = 0;
        if (drawn.x <= drawn.y) {
            s4 = make_int4(drawn.x, drawn.y, max(pg.px_lo, floor_div4(h.x_min)), min(pg.px_hi, floor_div4(h.x_max)));
            bands = drawn.y / kBandRows - drawn.x / kBandRows + 1;
        }
        sel[p] = s4;
        nb[p] = bands;
        if (feat) {
            // the dispatcher's 12 per-path features (sizes in device pixels)
            const rr_path_attr at = attrs[p];
            const float E = static_cast<float>(e1 - e0);
            const float W = s4.z <= s4.w ? static_cast<float>(s4.w - s4.z + 1) : 0.f;
            const float Hh = drawn.x <= drawn.y ? static_cast<float>(drawn.y - drawn.x + 1) : 0.f;
            const float area = W * Hh;
            float *f = feat + p * kFeatures;
            f[0] = log2f(1.f + E);
            f[1] = log2f(1.f + W);
            f[2] = log2f(1.f + Hh);
            f[3] = log2f(1.f + area);
            f[4] = log2f(1.f + E / fmaxf(1.f, 4.f * Hh));
            f[5] = static_cast<float>(bands);
            f[6] = (at.eo & RR_FILL_EVEN_ODD) ? 1.f : 0.f;
            f[7] = log2f((W + 1.f) / (Hh + 1.f));
            f[8] = fminf(1.f, static_cast<float>(fabs(0.5 * shoelace)) / fmaxf(1.f, area));
            f[9] = log2f(1.f + static_cast<float>(4.0 * dy_sum) / fmaxf(1.f, E));
            f[10] = static_cast<float>(at.a_input) / 255.f;
            f[11] = log2f(1.f + area / fmaxf(1.f, static_cast<float>(pg.width) * static_cast<float>(pg.height)));
        }
    }
}

__global__ void __launch_bounds__(256) k_edge_setup(const double *__restrict__ edges,
                                                    const int64_t *__restrict__ edge_off, int64_t n_paths,
                                                    const DevPage *__restrict__ pages, int32_t n_pages,
                                                    EdgeRec *__restrict__ recs, PathHdr *__restrict__ hdrs,
                                                    int2 *__restrict__ rows, int4 *__restrict__ sel,
                                                    int32_t *__restrict__ nb, const rr_path_attr *__restrict__ attrs,
                                                    float *__restrict__ feat, int2 *__restrict__ csum, int32_t mode,
                                                    int32_t *__restrict__ err) {
    __shared__ int64_t pb[kSetupPagesLds];
    const bool in_lds = n_pages <= kSetupPagesLds;
    if (in_lds)
        for (int32_t i = threadIdx.x; i < n_pages; i += blockDim.x) pb[i] = pages[i].path_begin;
    __syncthreads();
    const int64_t stride = static_cast<int64_t>(gridDim.x) * 4;
    const int32_t wave = __builtin_amdgcn_readfirstlane(static_cast<int32_t>(threadIdx.x >> 6));
    int32_t lo = 0;  // page of the current path (pages tile the paths in order)
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * 4 + wave; p < n_paths; p += stride) {
        if (in_lds) {
            while (lo + 1 < n_pages && pb[lo + 1] <= p) ++lo;
        } else {
            int32_t hi = n_pages - 1;
            while (lo < hi) {
                const int32_t mid = (lo + hi + 1) / 2;
                if (pages[mid].path_begin <= p)
                    lo = mid;
                else
                    hi = mid - 1;
            }
        }
        lo = __builtin_amdgcn_readfirstlane(lo);
        edge_setup_path(edges, edge_off, p, pages, lo, recs, hdrs, rows, sel, nb, attrs, feat, csum, mode, err);
    }
}

// Row summary of every 64-edge chunk (global edge index / 64): the smallest r0
// and largest r1 of its non-empty records, so that a band skips whole chunks.

// draw_aa_line's pixel count of the band (the reference composites x0..x1 of
// every drawn row, x0 = xx_min/4, x1 = (xx_max-1)/4 over its 4 AA rows,
// scanner/mod.rs:459-466, clamped by the clip): summed on the rows' first
// lanes.  X0/X1: each lane's disjoint sorted intervals ([k * 64 + lane]).
template <typename Ptr>
__device__ __forceinline__ unsigned long long band_draw_width(const DevPage &pg, int32_t Y0, int32_t ylo,
                                                              int32_t yhi, int nint, Ptr X0, Ptr X1) {
    const int lane = threadIdx.x & 63;
    const int32_t y = Y0 + (lane >> 2);
    int32_t smin = nint ? X0[lane] : INT32_MAX;
    int32_t emax = nint ? X1[(nint - 1) * 64 + lane] : -1;
    smin = min(smin, __shfl_xor(smin, 1));
    smin = min(smin, __shfl_xor(smin, 2));
    emax = max(emax, __shfl_xor(emax, 1));
    emax = max(emax, __shfl_xor(emax, 2));
    unsigned long long w = 0;
    if ((lane & 3) == 0 && y >= ylo && y <= yhi && smin < emax) {
        const int32_t x0 = max(smin / 4, pg.px_lo), x1 = min((emax - 1) / 4, pg.px_hi);
        if (x0 <= x1) w = static_cast<unsigned long long>(x1 - x0 + 1);
    }
    return w;
}

// Composite sweep of one band: this wave takes the CH-pixel steps wv, wv +
// nwv, ... of the band's covered columns, builds each step's 64-row coverage
// bitmap in its own `bm` (64 rows of CH/8 dwords, padded by one), and
// composites 4 pixels per lane per dword.  X0/X1: each lane's disjoint sorted
// intervals; any_unb: a row of the band never closed its winding (its span
// runs to the page edge).  Adds the path's coverage hash / pixel counts (and
// w_draw, the caller's draw_aa_line count) when asked.
template <int CH, bool WS, typename Ptr>
__device__ __forceinline__ void composite_band(const FillArgs &A, uint32_t *bm, const uint8_t *gamma,
                                               const DevPage &pg, int64_t p, int32_t Y0, const PathHdr &h,
                                               int32_t ylo, int32_t yhi, const rr_path_attr &at, int nint, Ptr X0,
                                               Ptr X1, bool any_unb, int wv, int nwv, unsigned long long w_draw) {
    const int lane = threadIdx.x & 63;
    // pixel range that can hold coverage, intersected with the clip
    const int32_t px_lo = max(pg.px_lo, floor_div4(h.x_min));
    const int32_t px_hi = min(pg.px_hi, any_unb ? pg.width - 1 : floor_div4(h.x_max));
    const bool any_px = px_lo <= px_hi;
    const int32_t xs = px_lo & ~3;
    const uint32_t a_in = at.a_input, src = at.color, fl = at.eo;
    // RR_FLAVOUR_RGB8: three channel planes of the path's colour; every other
    // flavour one grey byte a pixel
    const bool rgb = A.flavour == RR_FLAVOUR_RGB8, fast = A.flavour != RR_FLAVOUR_GRAY8;
    const uint32_t col = rgb ? A.colors[p] : src * 0x010101u;
    const bool opaque = a_in == 255u && !(fl & kFillGeneral);
    const bool use_mask = (fl & RR_FILL_SOFT_MASK) && pg.smask;
    uint64_t hsum = 0;
    unsigned long long w_cov = 0;
    int ip = 0;  // first interval of this lane not yet fully left of the chunk
    constexpr int kWords = CH / 8, kPitch = kWords + 1;  // bitmap dwords per AA row, padded
    constexpr int kLanesPerRow = CH / 4;            // one dword (4 pixels) per lane
    constexpr int kRowsPerIter = 64 / kLanesPerRow;  // output rows composited per pass
    constexpr int kIters = kBandRows / kRowsPerIter;
    constexpr int kGroup = kIters < 4 ? kIters : 4;
    uint32_t *bm_row = bm + lane * kPitch;
    const uint32_t src4 = src * 0x01010101u;
    const int q = lane % kLanesPerRow, sub = lane / kLanesPerRow;
    for (int32_t cx = xs + wv * CH; any_px && cx <= px_hi; cx += nwv * CH) {
        const int32_t A0 = 4 * cx, A1 = A0 + 4 * CH;
        while (ip < nint && X1[ip * 64 + lane] <= A0) ++ip;
        if (__ballot(ip < nint && X0[ip * 64 + lane] < A1) == 0) continue;  // nothing covered here
        for (int w = 0; w < kWords; ++w) bm_row[w] = 0u;
        for (int k = ip; k < nint; ++k) {
            const int32_t s = X0[k * 64 + lane];
            if (s >= A1) break;
            const int32_t lo = max(s, A0) - A0, hi = min(X1[k * 64 + lane], A1) - A0;
            const int wl = lo >> 5, wh = (hi - 1) >> 5;
            const uint32_t ml = ~0u << (lo & 31);
            const uint32_t mh = (hi & 31) ? ((1u << (hi & 31)) - 1u) : ~0u;
            if (wl == wh) {
                bm_row[wl] |= ml & mh;
            } else {
                bm_row[wl] |= ml;
                for (int w = wl + 1; w < wh; ++w) bm_row[w] = ~0u;
                bm_row[wh] |= mh;
            }
        }
        band_sync<WS>();
        const int32_t xq = cx + 4 * q;  // this lane's 4 pixels
        const int sh = (q & 1) * 16;
        const bool lane_in = xq <= px_hi;
        const bool full_ok = opaque && xq >= px_lo && xq + 3 <= px_hi;
        // Rows in groups: coverage from the bitmap, then the canvas loads of
        // the group issued back to back, then composite + store.
        for (int i0 = 0; i0 < kIters; i0 += kGroup) {
            uint32_t cov[kGroup], old[kGroup];
#pragma unroll
            for (int k = 0; k < kGroup; ++k) {
                const int dd = (i0 + k) * kRowsPerIter + sub;
                const int32_t yy = Y0 + dd;
                uint32_t cv = 0;
                if (lane_in && yy >= ylo && yy <= yhi) {
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        const uint32_t f = (bm[(4 * dd + rr) * kPitch + (q >> 1)] >> sh) & 0xffffu;
                        uint32_t t = f - ((f >> 1) & 0x5555u);
                        t = (t & 0x3333u) + ((t >> 2) & 0x3333u);
                        cv += (t & 0xfu) | ((t & 0xf0u) << 4) | ((t & 0xf00u) << 8) | ((t & 0xf000u) << 12);
                    }
                }
                cov[k] = cv;
                old[k] = 0;
                if (!rgb && cv != 0 && !(cv == 0x10101010u && full_ok))
                    old[k] = *reinterpret_cast<const uint32_t *>(pg.canvas + static_cast<long long>(yy) * pg.stride + xq);
            }
#pragma unroll
            for (int k = 0; k < kGroup; ++k) {
                const uint32_t cv = cov[k];
                if (cv == 0) continue;
                const int32_t yy = Y0 + (i0 + k) * kRowsPerIter + sub;
                const bool full = cv == 0x10101010u && full_ok;
                // soft-mask bytes of the 4 pixels (page space, 4-byte aligned rows)
                const uint32_t smw = use_mask && !full
                                         ? *reinterpret_cast<const uint32_t *>(pg.smask + static_cast<long long>(yy) *
                                                                                              pg.smask_stride + xq)
                                         : ~0u;
                if (rgb) {
                    uint32_t *dst = reinterpret_cast<uint32_t *>(pg.canvas + static_cast<long long>(yy) * pg.stride + 3 * xq);
                    uint32_t pl[3];
                    const uint32_t w0 = full ? 0u : dst[0], w1 = full ? 0u : dst[1], w2 = full ? 0u : dst[2];
                    if (full) {
#pragma unroll
                        for (int c = 0; c < 3; ++c) pl[c] = ((col >> (8 * c)) & 0xffu) * 0x01010101u;
                    } else {
                        deinterleave4(w0, w1, w2, pl[0], pl[1], pl[2]);
                        for (int px = 0; px < 4; ++px) {
                            const uint32_t cnt = (cv >> (8 * px)) & 0xffu;
                            const int32_t x = xq + px;
                            if (cnt == 0 || x < px_lo || x > px_hi) continue;
                            const uint32_t shape = gamma[cnt], sm = (smw >> (8 * px)) & 0xffu;
#pragma unroll
                            for (int c = 0; c < 3; ++c) {
                                const uint32_t res = paint_px((pl[c] >> (8 * px)) & 0xffu, shape, a_in,
                                                              (col >> (8 * c)) & 0xffu, fl, sm, fast);
                                pl[c] = (pl[c] & ~(0xffu << (8 * px))) | (res << (8 * px));
                            }
                        }
                    }
                    uint32_t o0, o1, o2;
                    interleave4(pl[0], pl[1], pl[2], o0, o1, o2);
                    if (full || o0 != w0) dst[0] = o0;
                    if (full || o1 != w1) dst[1] = o1;
                    if (full || o2 != w2) dst[2] = o2;
                } else {
                    uint32_t *dst = reinterpret_cast<uint32_t *>(pg.canvas + static_cast<long long>(yy) * pg.stride + xq);
                    uint32_t out;
                    if (full) {
                        out = src4;  // full coverage of an opaque source: both composites give src
                    } else {
                        out = old[k];
                        for (int px = 0; px < 4; ++px) {
                            const uint32_t cnt = (cv >> (8 * px)) & 0xffu;
                            const int32_t x = xq + px;
                            if (cnt == 0 || x < px_lo || x > px_hi) continue;
                            const uint32_t res = paint_px((out >> (8 * px)) & 0xffu, gamma[cnt], a_in, src, fl,
                                                          (smw >> (8 * px)) & 0xffu, fast);
                            out = (out & ~(0xffu << (8 * px))) | (res << (8 * px));
                        }
                    }
                    if (out != old[k] || full) *dst = out;
                }
                if (A.cov_count || A.cov_hash) {
                    for (int px = 0; px < 4; ++px) {
                        const uint32_t cnt = (cv >> (8 * px)) & 0xffu;
                        const int32_t x = xq + px;
                        if (cnt == 0 || x < px_lo || x > px_hi) continue;
                        ++w_cov;
                        if (A.cov_hash) hsum += cov_term(static_cast<long long>(yy) * pg.width + x, gamma[cnt]);
                    }
                }
            }
        }
        band_sync<WS>();
    }
    if (A.cov_hash) {
        for (int o = 32; o > 0; o >>= 1) hsum += __shfl_xor(hsum, o);
        if (lane == 0 && hsum) atomicAdd(reinterpret_cast<unsigned long long *>(A.cov_hash + p), hsum);
    }
    if (A.cov_count) {
        for (int o = 32; o > 0; o >>= 1) {
            w_cov += __shfl_xor(w_cov, o);
            w_draw += __shfl_xor(w_draw, o);
        }
        if (lane == 0) {
            if (w_cov) atomicAdd(reinterpret_cast<unsigned long long *>(A.cov_count + 2 * p), w_cov);
            if (w_draw) atomicAdd(reinterpret_cast<unsigned long long *>(A.cov_count + 2 * p + 1), w_draw);
        }
    }
}

// Rasterises path p over output rows [Y0, Y0+15] of page `pg` and composites
// into its canvas.  Called by one full wave64 (lane = threadIdx.x & 63): the
// whole workgroup, or with WS one wave of a larger one working alone; `slot`
// is this wave's private overflow slot.
template <int K, int CH, bool WS>
__device__ __forceinline__ void fill_band(const FillArgs &A, BandLds<K, CH> &L, const DevPage &pg, int64_t p, int32_t Y0,
                          int64_t slot) {
    const int lane = threadIdx.x & 63;
    // independent loads first, so their latencies overlap
    const PathHdr h = A.hdrs[p];
    const int2 drawn = A.rows[p];
    const rr_path_attr at = A.attrs[p];
    const int64_t e0 = A.edge_off[p], e1 = A.edge_off[p + 1];
    if (h.flags & (HDR_EMPTY | HDR_OUTSIDE)) return;
    const int32_t ylo = drawn.x, yhi = drawn.y;
    if (Y0 > yhi || Y0 + kBandRows - 1 < ylo) return;
    const bool literal = A.mode == RR_MODE_LITERAL;

    const int32_t d = lane >> 2, j = lane & 3;
    const int32_t y = Y0 + d;
    const int32_t r = literal ? 16 * y + 12 + j : 4 * y + j;
    const int32_t band_lo = literal ? 16 * Y0 + 12 : 4 * Y0;
    const int32_t band_hi = literal ? 16 * (Y0 + kBandRows - 1) + 15 : 4 * Y0 + 4 * kBandRows - 1;
    const bool valid = y >= ylo && y <= yhi && r >= h.y_min && r <= h.y_max;
    const int32_t mrule = (at.eo & RR_FILL_EVEN_ODD) ? 1 : -1;
    const int32_t aa_width = pg.width * 4;

    int32_t *X0 = L.ex0, *X1 = L.ex1, *C = L.ec;
    const int2 sm0 = chunk_summary_first(A, e0, e1);
    int n = build_entries<WS>(A, L, e0, e1, band_lo, band_hi, literal, Y0, valid, r, mrule, X0, X1, C, K, sm0);
    int need = n;
    for (int o = 32; o > 0; o >>= 1) need = max(need, __shfl_xor(need, o));
    if (need > K) {
        // A row with more state-changing entries than the LDS holds: redo the
        // band with the entry lists in this wave's global slot, or for very
        // long rows in a slice of the shared pool.
        int32_t *G;
        if (static_cast<long long>(need) * 192 <= A.slot_ints) {
            G = A.slots + slot * A.slot_ints;
        } else {
            unsigned long long off = 0;
            if (lane == 0) off = atomicAdd(A.spill_top, static_cast<unsigned long long>(need) * 192ull);
            off = __shfl(off, 0);
            if (static_cast<long long>(off) + static_cast<long long>(need) * 192 > A.spill_cap) {
                if (lane == 0) atomicExch(A.err, 1);
                return;
            }
            G = A.spill + off;
        }
        X0 = G;
        X1 = G + static_cast<long long>(need) * 64;
        C = G + static_cast<long long>(need) * 128;
        n = build_entries<WS>(A, L, e0, e1, band_lo, band_hi, literal, Y0, valid, r, mrule, X0, X1, C, need, sm0);
    }
    bool unb = false;
    const int nint = sort_and_walk(n, X0, X1, C, mrule, aa_width, &unb);
    const bool any_unb = __ballot(unb) != 0;
    const unsigned long long w_draw = A.cov_count ? band_draw_width(pg, Y0, ylo, yhi, nint, X0, X1) : 0ull;
    composite_band<CH, WS>(A, L.bm, L.gamma, pg, p, Y0, h, ylo, yhi, at, nint, X0, X1, any_unb, 0, 1, w_draw);
}

// ---------------------------------------------------------------------------
// PLAIN: several waves per band
// ---------------------------------------------------------------------------
// A one-path launch holds only that path's bands, so a band's time is the
// latency of its serial chain.  Here a group of NW waves shares the band.
// Pass by pass, wave w of the group takes the w-th of the next NW edge chunks
// (64 edges each) that overlap the band, marks per AA row (one lane per row)
// the chunk edges crossing it, and writes each crossing's intersection
// (fill_intersections, scanner/mod.rs:528-585) into the row's LDS list at its
// place in edge order: the waves' crossing counts are summed in wave order
// ahead of the writes, so the list comes out in the order the reference walks
// the edges.  The group's first wave then folds each row's list with
// write_intersect's merge (scanner/mod.rs:600-642) -- a short loop over LDS
// instead of a loop over the path's edges.  The stable sort by x0 is a rank
// count split over the group's waves, the walk is render_aa_line's, and the
// composite sweep's CH-pixel steps go to the waves in turn.
//
// A workgroup holds one group (one band; the launch has a workgroup per band)
// or G groups (the path's G bands in one workgroup: the launch is a single
// workgroup, which the device dispatches at a higher rate than several).  The
// groups of a workgroup pass the same barriers: every group runs as many chunk
// passes as the busiest one.  Intersections are kept packed in 4 bytes, x0 and
// x1 as 15-bit offsets from the path's leftmost AA column plus the 2-bit count,
// so that three bands' lists fit one workgroup's LDS.  A row with more
// intersections or entries than the LDS lists hold, or with an x beyond the
// 15-bit range, sends the band, untouched, to the one-wave fill_band (the
// group's first wave, global entry slots) once the groups' shared barriers are
// behind.
constexpr int kMwWaves = 16;  // waves per band, one band per workgroup
constexpr int kMwPool = 128;  // intersections per AA row of a band in LDS
constexpr int kMwK = 40;      // entries per AA row after the merge
constexpr int kMwBlock = 8;   // entries read into registers at a time
constexpr uint32_t kMwSpan = 32767u;  // largest packed x offset (AA columns)

template <int NW, int P, int K2, int CH>
struct MwLds {
    static constexpr int kPitch = CH / 8 + 1;
    struct Build {
        int4 recA[NW][64], recB[NW][64];   // each wave's chunk as two 16-byte halves per edge
        unsigned long long smask[NW][64];  // per wave, per lane row: chunk edges crossing it
        int32_t cnt[2][NW][64];            // crossings of (wave's chunk, row), by pass parity
    };
    struct Merged {
        // write_intersect's entries [k * 64 + lane]: x0 | x1 << 16 as offsets
        // from the path's leftmost column, and the count; row K2 takes the
        // stores the branch-free loop discards
        uint32_t mx[(K2 + 1) * 64];
        int32_t mc[(K2 + 1) * 64];
    };
    union {
        Build b;
        Merged m;
        uint32_t bm[NW][64 * kPitch];  // each wave's coverage bitmap (composite sweep)
    } u;
    union {
        // the intersections of each row in edge order [k * 64 + lane],
        // (x0 - xb) | (x1 - xb) << 15 | (count + 1) << 30; row P takes the
        // stores the branch-free loop discards
        uint32_t pk[(P + 1) * 64];
        // then the merged entries sorted by x0 (sx packed like mx, sc the
        // count), and in place over them the row's disjoint intervals
        // (absolute AA columns, starts in sx, ends in sc); row K2 = discards
        struct {
            int32_t sx[(K2 + 1) * 64], sc[(K2 + 1) * 64];
        } s;
    } q;
    int32_t nb[64];
    uint8_t gamma[32];
    int32_t over, unb;
};

// The band at rows Y0.. of path p, filled and composited by the NW waves of
// group `grp` (of `ng` in the workgroup; npass: 2 x ng ints of LDS shared by
// the groups).  Returns false, with nothing drawn, when a row outgrew the LDS
// lists; every wave of the workgroup passes the same barriers either way.
template <int NW, int P, int K2, int CH>
__device__ __forceinline__ bool fill_band_mw(const FillArgs &A, MwLds<NW, P, K2, CH> &L, int32_t *npass, int ng,
                                             int grp, const DevPage &pg, int64_t p, int32_t Y0) {
    static_assert(2 * (K2 + 1) <= P + 1, "the sorted entries live in the intersection lists");
    static_assert(K2 < 256, "sort keys hold the entry index in 8 bits");
    const int wv = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)) - grp * NW;
    const int lane = threadIdx.x & 63;
    const int lt = wv * 64 + lane;  // thread index within the group
    const PathHdr h = A.hdrs[p];
    const int2 drawn = A.rows[p];
    const rr_path_attr at = A.attrs[p];
    const int64_t e0 = A.edge_off[p], e1 = A.edge_off[p + 1];
    if (h.flags & (HDR_EMPTY | HDR_OUTSIDE)) return true;  // the same for every group
    const int32_t ylo = drawn.x, yhi = drawn.y;
    // a band outside the drawn rows has nothing to do but the barriers
    const bool active = !(Y0 > yhi || Y0 + kBandRows - 1 < ylo);
    const bool literal = A.mode == RR_MODE_LITERAL;
    const int32_t y = Y0 + (lane >> 2);
    const int32_t r = literal ? 16 * y + 12 + (lane & 3) : 4 * y + (lane & 3);
    const int32_t band_lo = literal ? 16 * Y0 + 12 : 4 * Y0;
    const int32_t band_hi = literal ? 16 * (Y0 + kBandRows - 1) + 15 : 4 * Y0 + 4 * kBandRows - 1;
    const bool valid = active && y >= ylo && y <= yhi && r >= h.y_min && r <= h.y_max;
    const int32_t mrule = (at.eo & RR_FILL_EVEN_ODD) ? 1 : -1;
    // every intersection lies in [h.x_min, h.x_max]; x is kept as x - xb
    const int32_t xb = h.x_min;
    if (lt < 17) L.gamma[lt] = kGamma[lt];
    if (lt == 0) L.over = 0;
    __syncthreads();

    int top = 0;  // intersections of this lane's row so far (every wave keeps the same count)
    int par = 0;
    bool wide = false;  // an x beyond the packed range
    if (e1 > e0) {
        const int64_t c_first = e0 >> 6, c_last = (e1 - 1) >> 6;
        int np_par = 0;
        for (int64_t cb = c_first; cb <= c_last; cb += 64) {
            bool hit = false;
            if (active && cb + lane <= c_last) {
                const int2 sm = A.csum[cb + lane];
                hit = sm.x <= band_hi && sm.y >= band_lo;
            }
            uint64_t cm = __ballot(hit);  // the same in every wave of the group
            int passes = (__popcll(cm) + NW - 1) / NW;
            if (ng > 1) {
                // as many passes as the busiest group (double-buffered: a
                // group writes a slot again only after the next barrier)
                if (wv == 0 && lane == 0) npass[np_par * ng + grp] = passes;
                __syncthreads();
                for (int g = 0; g < ng; ++g) passes = max(passes, npass[np_par * ng + g]);
                np_par ^= 1;
            }
            // this wave's chunk of a pass: the wv-th set bit of the mask left
            auto pick = [&](uint64_t m) {
                for (int i = 0; i < wv; ++i) m &= m - 1;
                return m;
            };
            auto load_rec = [&](uint64_t m) {
                EdgeRec q{1, 0, 0, 0, 0, 0, 0, 0};
                const int64_t i = ((cb + (__ffsll(static_cast<unsigned long long>(m)) - 1)) << 6) + lane;
                if (i >= e0 && i < e1) q = A.recs[i];
                return q;
            };
            uint64_t mine = pick(cm);
            EdgeRec qn{1, 0, 0, 0, 0, 0, 0, 0};
            if (mine) qn = load_rec(mine);
            for (int it = 0; it < passes; ++it) {
                for (int i = 0; i < NW; ++i) cm &= cm - 1;
                const EdgeRec q = qn;
                const bool have = mine != 0;
                mine = pick(cm);
                if (mine) qn = load_rec(mine);  // the next pass's chunk, in flight during this one
                unsigned long long m = 0ull;
                if (have) {
                    L.u.b.recA[wv][lane] = make_int4(q.r0, q.r1, q.meta, q.a);
                    L.u.b.recB[wv][lane] = make_int4(q.b, q.c, q.d, q.e);
                    int la = 64, lb = -1;
                    if (q.r0 <= q.r1) lane_range(q.r0, q.r1, literal, Y0, band_lo, band_hi, &la, &lb);
                    L.u.b.smask[wv][lane] = 0ull;
                    wave_sync();
                    // edge j toggles its bit on at lane la and off at lane lb + 1
                    const unsigned long long from0 = __ballot(la == 0 && lb >= 0);
                    if (la <= lb) {
                        if (la > 0) atomicXor(&L.u.b.smask[wv][la], 1ull << lane);
                        if (lb < 63) atomicXor(&L.u.b.smask[wv][lb + 1], 1ull << lane);
                    }
                    wave_sync();
                    m = wave_incl_xor64(L.u.b.smask[wv][lane]) ^ from0;
                    if (!valid) m = 0ull;
                }
                L.u.b.cnt[par][wv][lane] = __popcll(m);
                __syncthreads();
                // this wave's place in the row's list: after the chunks before it
                int off = top;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const int c = L.u.b.cnt[par][w][lane];
                    off += w < wv ? c : 0;
                    top += c;
                }
                par ^= 1;
                // the intersections, in edge order; the next record's LDS
                // load is issued ahead
                int4 ga = make_int4(0, 0, 0, 0), gb = ga;
                if (m) {
                    const int jj = __ffsll(m) - 1;
                    ga = L.u.b.recA[wv][jj];
                    gb = L.u.b.recB[wv][jj];
                }
                while (m) {
                    m &= m - 1;
                    const EdgeRec g{ga.x, ga.y, ga.z, ga.w, gb.x, gb.y, gb.z, gb.w};
                    if (m) {
                        const int jn = __ffsll(m) - 1;
                        ga = L.u.b.recA[wv][jn];
                        gb = L.u.b.recB[wv][jn];
                    }
                    const Entry en = row_entry(g, r);
                    const uint32_t a0 = static_cast<uint32_t>(en.x0) - static_cast<uint32_t>(xb);
                    const uint32_t a1 = static_cast<uint32_t>(en.x1) - static_cast<uint32_t>(xb);
                    wide = wide || a0 > kMwSpan || a1 > kMwSpan;
                    const int slot = min(off, P);
                    L.q.pk[slot * 64 + lane] = (a0 & kMwSpan) | (a1 & kMwSpan) << 15 |
                                               static_cast<uint32_t>(en.c + 1) << 30;
                    ++off;
                }
            }
        }
    }
    if (__ballot(top > P || wide) != 0 && lane == 0) L.over = 1;
    __syncthreads();
    // Entries are read from LDS kMwBlock at a time into registers and the
    // loops over them are unrolled and branch-free (a store that is not
    // wanted goes to the spare row).
    constexpr int kB = kMwBlock;
    if (wv == 0 && L.over == 0) {
        // write_intersect over the row's intersections: merge into the open
        // entry while they touch it; a retired entry is kept only when it can
        // change the walk's inside/outside state
        int32_t lx0 = INT32_MAX, lx1 = INT32_MIN, lc = 0;
        int nb = 0;
        for (int k0 = 0; k0 < top; k0 += kB) {
            int32_t x0[kB], x1[kB], cc[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const uint32_t v = L.q.pk[min(k0 + j, P) * 64 + lane];
                x0[j] = static_cast<int32_t>(v & kMwSpan);
                x1[j] = static_cast<int32_t>((v >> 15) & kMwSpan);
                cc[j] = static_cast<int32_t>(v >> 30) - 1;
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const bool v = k0 + j < top;
                const bool touch = lx1 + 1 >= x0[j] && lx0 <= x1[j] + 1;
                const bool st = v && !touch && (lc & mrule) != 0;
                const int slot = st && nb < K2 ? nb : K2;
                L.u.m.mx[slot * 64 + lane] = static_cast<uint32_t>(lx0) | static_cast<uint32_t>(lx1) << 16;
                L.u.m.mc[slot * 64 + lane] = lc;
                nb += st ? 1 : 0;
                const int32_t sum = static_cast<int32_t>(static_cast<uint32_t>(lc) + static_cast<uint32_t>(cc[j]));
                lc = v ? (touch ? sum : cc[j]) : lc;
                lx0 = v ? (touch ? min(lx0, x0[j]) : x0[j]) : lx0;
                lx1 = v ? (touch ? max(lx1, x1[j]) : x1[j]) : lx1;
            }
        }
        if ((lc & mrule) != 0) {
            const int slot = nb < K2 ? nb : K2;
            L.u.m.mx[slot * 64 + lane] = static_cast<uint32_t>(lx0) | static_cast<uint32_t>(lx1) << 16;
            L.u.m.mc[slot * 64 + lane] = lc;
            ++nb;
        }
        L.nb[lane] = min(nb, K2);
        if (__ballot(nb > K2) != 0 && lane == 0) L.over = 1;
    }
    __syncthreads();
    const bool fallback = L.over != 0;  // written only ahead of the barrier above

    // stable sort by x0: entry i of the row goes to its rank; each thread
    // ranks the entries wv, wv + NW, ... against every entry of the row
    if (!fallback) {
        // key = x0 * 256 + index: one 64-bit compare per pair; entries past
        // the row's end get the largest key
        constexpr int kPer = (K2 + NW - 1) / NW;
        const int n = L.nb[lane];
        long long key[kPer];
        int rank[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int i = wv + e * NW;
            const uint32_t x = L.u.m.mx[min(i, K2) * 64 + lane] & 0xffffu;
            key[e] = i < n ? static_cast<long long>(x) * 256 + i : LLONG_MAX;
            rank[e] = 0;
        }
        for (int k0 = 0; k0 < n; k0 += kB) {
            long long kk[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const int k = k0 + j;
                const uint32_t x = L.u.m.mx[min(k, K2) * 64 + lane] & 0xffffu;
                kk[j] = k < n ? static_cast<long long>(x) * 256 + k : LLONG_MAX;
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
#pragma unroll
                for (int e = 0; e < kPer; ++e) rank[e] += kk[j] < key[e] ? 1 : 0;
            }
        }
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const int i = wv + e * NW;
            if (i < n) {
                L.q.s.sx[rank[e] * 64 + lane] = static_cast<int32_t>(L.u.m.mx[i * 64 + lane]);
                L.q.s.sc[rank[e] * 64 + lane] = L.u.m.mc[i * 64 + lane];
            }
        }
    }
    __syncthreads();
    unsigned long long w_draw = 0;
    if (wv == 0 && !fallback) {
        // render_aa_line's walk (scanner/mod.rs:393-467) over the sorted
        // entries: an entry met outside opens a span at its x0, the entry that
        // brings the winding back to outside closes it at its x1 + 1; spans
        // clamped to [0, aa_width) and merged into disjoint intervals, written
        // in place over the entries already read
        const int n = L.nb[lane];
        const int32_t aa_width = pg.width * 4;
        int32_t count = 0, xx = 0, cs = 0, ce = -1;
        bool open = false;
        int nint = 0;
        auto put = [&](int32_t bx0, int32_t bx1, bool go) {
            // merge [bx0, bx1) into the current interval or start a new one
            const bool any = go && bx0 < bx1;
            const bool join = ce >= 0 && bx0 <= ce;
            const bool emit = any && !join && ce >= 0;
            const int slot = emit ? nint : K2;
            L.q.s.sx[slot * 64 + lane] = cs;
            L.q.s.sc[slot * 64 + lane] = ce;
            nint += emit ? 1 : 0;
            cs = any && !join ? bx0 : cs;
            ce = any ? (join ? max(ce, bx1) : bx1) : ce;
        };
        for (int k0 = 0; k0 < n; k0 += kB) {
            int32_t x0[kB], x1[kB], cc[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                // an interval is written at most at the index of the entry
                // that closes it, so this block is read before any write to it
                const int k = min(k0 + j, K2);
                const uint32_t v = static_cast<uint32_t>(L.q.s.sx[k * 64 + lane]);
                x0[j] = static_cast<int32_t>(static_cast<uint32_t>(xb) + (v & 0xffffu));
                x1[j] = static_cast<int32_t>(static_cast<uint32_t>(xb) + (v >> 16));
                cc[j] = L.q.s.sc[k * 64 + lane];
            }
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                const bool v = k0 + j < n;
                count = v ? static_cast<int32_t>(static_cast<uint32_t>(count) + static_cast<uint32_t>(cc[j])) : count;
                const bool opener = v && !open;
                const bool closer = v && open && (count & mrule) == 0;
                xx = opener ? x0[j] : xx;
                const int32_t end = x1[j] == INT32_MAX ? INT32_MAX : x1[j] + 1;
                put(max(xx, 0), max(min(end, aa_width), 0), closer);
                open = opener ? true : (closer ? false : open);
            }
        }
        // a span the winding never closed runs to the row's end
        const bool unb = open;
        put(max(xx, 0), aa_width, unb);
        if (ce >= 0) {
            L.q.s.sx[nint * 64 + lane] = cs;
            L.q.s.sc[nint * 64 + lane] = ce;
            ++nint;
        }
        const bool any_unb = __ballot(unb) != 0;
        L.nb[lane] = nint;
        if (lane == 0) L.unb = any_unb ? 1 : 0;
        if (A.cov_count) w_draw = band_draw_width(pg, Y0, ylo, yhi, nint, L.q.s.sx, L.q.s.sc);
    }
    __syncthreads();
    if (fallback) return false;
    composite_band<CH, true>(A, L.u.bm[wv], L.gamma, pg, p, Y0, h, ylo, yhi, at, L.nb[lane], L.q.s.sx, L.q.s.sc,
                             L.unb != 0, wv, NW, w_draw);
    return true;
}

template <int NW, int P, int K2, int CH, int FK>
union MwUnion {
    MwLds<NW, P, K2, CH> m;
    BandLds<FK, CH> f;  // the one-wave fallback
};

// Writes the plain launches' shared arguments to device memory (launch
// arguments are copied when the launch is issued, so the host may reuse A at
// once).
__global__ void __launch_bounds__(64) k_store_fill_args(FillArgs A, FillArgs *dst) {
    if (threadIdx.x == 0) *dst = A;
}

// PLAIN: one launch per path.  G == 1: a workgroup of NW waves per band of 16
// output rows (grid = the path's bands); G > 1: one workgroup whose groups of
// NW waves take the path's bands (blockDim = 64 * NW * bands, bands <= G).
// The recording's shared arguments reach the launches through device memory
// (Ap, written by k_store_fill_args ahead of them): a launch carries 32 bytes
// of arguments instead of the 160-byte FillArgs, less for the host to copy on
// each of the thousands of launches of a page.
template <int G, int NW, int P, int K2, int CH, int FK>
__global__ void __launch_bounds__(64 * G * NW) k_fill_plain(const FillArgs *__restrict__ Ap, int32_t page, int64_t p,
                                                            int32_t row0, int32_t slot0) {
    static_assert(sizeof(MwUnion<NW, P, K2, CH, FK>) * G + 8 * G <= 163840, "LDS of one workgroup");
    const FillArgs &A = *Ap;
    __shared__ MwUnion<NW, P, K2, CH, FK> U[G];
    __shared__ int32_t npass[2 * G];
    const int grp = G == 1 ? 0 : __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)) / NW;
    const int ng = G == 1 ? 1 : static_cast<int>(blockDim.x) / (64 * NW);
    const int band = G == 1 ? static_cast<int>(blockIdx.x) : grp;
    const DevPage &pg = A.pages[page];
    const int32_t Y0 = row0 + kBandRows * band;
    const int lt = static_cast<int>(threadIdx.x) - grp * NW * 64;
    const long long t0 = A.trace ? static_cast<long long>(wall_clock64()) : 0;
    if (!fill_band_mw<NW, P, K2, CH>(A, U[grp].m, npass, ng, grp, pg, p, Y0)) {
        // past the groups' last shared barrier: the group's first wave alone
        if (lt >= 64) return;
        if (lt < 17) U[grp].f.gamma[lt] = kGamma[lt];
        wave_sync();
        fill_band<FK, CH, true>(A, U[grp].f, pg, p, Y0, slot0 + static_cast<int64_t>(band));
    }
    if (A.trace && lt == 0) {
        // wall-clock span of the group's first wave's share of the band
        const long long t1 = static_cast<long long>(wall_clock64());
        const unsigned long long k = atomicAdd(A.trace_top, 1ull);
        if (static_cast<long long>(k) < A.trace_cap)
            A.trace[k] = make_longlong4(p, (static_cast<long long>(page) << 32) | static_cast<uint32_t>(Y0), t0, t1);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Plain-launch kernels of every instantiated fill shape, by launch shape:
// [0] a workgroup of kMwWaves waves per band, [1] one workgroup for a path of
// two bands (8 waves each), [2] one workgroup for a path of three bands (5
// waves each).  CH from kVariantCH, at most 64 pixels a step on this route
// (which keeps the waves' bitmaps inside the LDS the intersection lists
// leave); FK = kVariantK, the LDS entry slots of the one-wave fallback.
constexpr int kPlainShapes = 3;
constexpr int kPlainGroups[kPlainShapes] = {1, 2, 3};
constexpr int kPlainGroupWaves[kPlainShapes] = {kMwWaves, 8, 5};
constexpr int plain_ch(int v) { return kVariantCH[v] > 64 ? 64 : kVariantCH[v]; }
#define RR_PLAIN(s, v)                                                                                    \
    reinterpret_cast<const void *>(&k_fill_plain<kPlainGroups[s], kPlainGroupWaves[s], kMwPool, kMwK, \
                                                 plain_ch(v), kVariantK[v]>)
#define RR_PLAIN_ROW(s)                                                                                   \
    {RR_PLAIN(s, 0),  RR_PLAIN(s, 1),  RR_PLAIN(s, 2),  RR_PLAIN(s, 3),  RR_PLAIN(s, 4),  RR_PLAIN(s, 5), \
     RR_PLAIN(s, 6),  RR_PLAIN(s, 7),  RR_PLAIN(s, 8),  RR_PLAIN(s, 9),  RR_PLAIN(s, 10), RR_PLAIN(s, 11), \
     RR_PLAIN(s, 12), RR_PLAIN(s, 13), RR_PLAIN(s, 14), RR_PLAIN(s, 15)}
const void *const kPlainKernels[kPlainShapes][kNumVariants] = {RR_PLAIN_ROW(0), RR_PLAIN_ROW(1), RR_PLAIN_ROW(2)};
#undef RR_PLAIN_ROW
#undef RR_PLAIN

}  // namespace

namespace {

constexpr long long kSlotInts = static_cast<long long>(kSlotK) * 192;

int32_t fail(rr_hip_ctx *c, int32_t code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int32_t hip_fail(rr_hip_ctx *c, hipError_t e, const char *what) {
    return fail(c, RR_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define RR_HIP_TRY(ctx, expr)                                  \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)


// Next free start/stop event pair of the per-launch profiling pool.
hipEvent_t *prof_pair(rr_hip_ctx *c) {
    if (c->prof_used + 2 > c->prof_ev.size()) {
        const size_t grow = c->prof_ev.size() ? c->prof_ev.size() : 256;
        for (size_t k = 0; k < grow; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c->prof_ev.push_back(e);
        }
    }
    hipEvent_t *p = &c->prof_ev[c->prof_used];
    c->prof_used += 2;
    return p;
}

int32_t sat_i32(double v) {
    if (std::isnan(v)) return INT32_MIN;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return static_cast<int32_t>(v);
}

// Clip::new + recompute_int_bounds (crates/raster/src/clip.rs:73-88,266-271)
ClipRect make_clip(const double cl[4]) {
    ClipRect r;
    r.x_min = std::fmin(cl[0], cl[2]);
    r.x_max = std::fmax(cl[0], cl[2]);
    r.y_min = std::fmin(cl[1], cl[3]);
    r.y_max = std::fmax(cl[1], cl[3]);
    r.x_min_i = sat_i32(std::floor(r.x_min));
    r.y_min_i = sat_i32(std::floor(r.y_min));
    r.x_max_i = sat_i32(std::ceil(r.x_max)) - 1;
    r.y_max_i = sat_i32(std::ceil(r.y_max)) - 1;
    return r;
}

int32_t validate_page(rr_hip_ctx *c, const rr_page_desc &p, int32_t flavour) {
    const int32_t bpp = flavour == RR_FLAVOUR_RGB8 ? 3 : 1;
    const int32_t w4 = (p.width + 3) & ~3;
    if (p.width <= 0 || p.height <= 0 || p.stride < bpp * w4 || (p.stride & 3) != 0)
        return fail(c, RR_E_INVARIANT,
                    "page: bad canvas geometry (stride must be a multiple of 4 >= bytes of width rounded up to 4)");
    if (!p.canvas || (reinterpret_cast<uintptr_t>(p.canvas) & 3u))
        return fail(c, RR_E_INVARIANT, "page: canvas must be a 4-byte aligned device pointer");
    if (p.path_begin < 0 || p.path_end < p.path_begin) return fail(c, RR_E_INVARIANT, "page: bad path range");
    if (p.soft_mask && ((reinterpret_cast<uintptr_t>(p.soft_mask) & 3u) || (p.soft_mask_stride & 3) ||
                        p.soft_mask_stride < w4))
        return fail(c, RR_E_INVARIANT, "page: soft mask must be 4-byte aligned with a stride a multiple of 4 >= width");
    if (p.gray && flavour != RR_FLAVOUR_RGB8) return fail(c, RR_E_INVARIANT, "page: a gray output needs RR_FLAVOUR_RGB8");
    if (p.gray && ((reinterpret_cast<uintptr_t>(p.gray) & 3u) || (p.gray_stride & 3) || p.gray_stride < w4))
        return fail(c, RR_E_INVARIANT, "page: gray output must be 4-byte aligned with a stride a multiple of 4 >= width");
    return RR_OK;
}

// Records the edge-setup pass and the fill kernels for every page of b.
int32_t record_pages(rr_hip_ctx *c, const rr_fill_pages &b) {
    if (!c->page_open) return fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (b.n_paths < 0 || b.n_edges < 0 || b.n_pages < 0) return fail(c, RR_E_INVARIANT, "negative counts");
    if (b.mode != RR_MODE_DOCUMENTED && b.mode != RR_MODE_LITERAL) return fail(c, RR_E_INVARIANT, "unknown AA row mode");
    if (b.flavour != RR_FLAVOUR_GRAY8 && b.flavour != RR_FLAVOUR_RGB8_FAST && b.flavour != RR_FLAVOUR_RGB8)
        return fail(c, RR_E_INVARIANT, "unknown canvas flavour");
    if (b.flavour == RR_FLAVOUR_RGB8 && b.n_paths > 0 && (!b.colors || (reinterpret_cast<uintptr_t>(b.colors) & 3u)))
        return fail(c, RR_E_INVARIANT, "RR_FLAVOUR_RGB8 needs a 4-byte aligned device array of path colours");
    if (b.launch != RR_LAUNCH_PLAIN && b.launch != RR_LAUNCH_BATCHED)
        return fail(c, RR_E_INVARIANT, "unknown launch shape");
    if (b.variant < 0 || b.variant >= kNumVariants) return fail(c, RR_E_INVARIANT, "unknown fill variant");
    if (b.n_paths == 0 || b.n_pages == 0) return RR_OK;
    if (!b.edges || !b.edge_off || !b.attrs || !b.pages || (b.launch == RR_LAUNCH_PLAIN && !b.host_rows))
        return fail(c, RR_E_INVARIANT, "null buffer");
    if (reinterpret_cast<uintptr_t>(b.edges) & 31u) return fail(c, RR_E_INVARIANT, "edges must be 32-byte aligned");
    if (reinterpret_cast<uintptr_t>(b.attrs) & 3u) return fail(c, RR_E_INVARIANT, "attrs must be 4-byte aligned");
    int64_t prev_end = 0;
    c->host_pages.resize(static_cast<size_t>(b.n_pages));
    for (int32_t i = 0; i < b.n_pages; ++i) {
        const rr_page_desc &p = b.pages[i];
        int32_t rc = validate_page(c, p, b.flavour);
        if (rc) return rc;
        if (p.path_begin != prev_end) return fail(c, RR_E_INVARIANT, "pages must tile [0, n_paths) in order");
        prev_end = p.path_end;
        DevPage &d = c->host_pages[static_cast<size_t>(i)];
        d.canvas = p.canvas;
        d.path_begin = p.path_begin;
        d.path_end = p.path_end;
        d.clip = make_clip(p.clip);
        d.width = p.width;
        d.height = p.height;
        d.stride = p.stride;
        d.px_lo = std::max(0, d.clip.x_min_i);
        d.px_hi = std::min(p.width - 1, d.clip.x_max_i);
        d.band_base = 0;
        d.cbase = 0;
        d.ctiles = 0;
        d.pad2 = 0;
        d.smask = p.soft_mask;
        d.smask_stride = p.soft_mask_stride;
        d.gray = p.gray;
        d.gray_stride = p.gray_stride;
    }
    if (prev_end != b.n_paths) return fail(c, RR_E_INVARIANT, "pages must tile [0, n_paths) in order");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    int32_t rc = rrint::ensure(c, &c->recs, &c->recs_cap, static_cast<size_t>(b.n_edges > 0 ? b.n_edges : 1));
    if (!rc) rc = rrint::ensure(c, &c->hdrs, &c->hdrs_cap, static_cast<size_t>(b.n_paths));
    if (!rc) rc = rrint::ensure(c, &c->rows, &c->rows_cap, static_cast<size_t>(b.n_paths));
    if (!rc) rc = rrint::ensure(c, &c->sel, &c->sel_cap, static_cast<size_t>(b.n_paths));
    if (!rc) rc = rrint::ensure(c, &c->nb, &c->nb_cap, static_cast<size_t>(b.n_paths));
    if (!rc && b.launch == RR_LAUNCH_BATCHED)
        rc = rrint::ensure(c, &c->feat, &c->feat_cap, static_cast<size_t>(b.n_paths) * kFeatures);
    if (!rc) rc = rrint::ensure(c, &c->pages, &c->pages_cap, static_cast<size_t>(b.n_pages));
    const int64_t n_chunks = (b.n_edges + 63) / 64;
    if (!rc) rc = rrint::ensure(c, &c->csum, &c->csum_cap, static_cast<size_t>(n_chunks > 0 ? n_chunks : 1));
    if (rc) return rc;
    if (b.launch == RR_LAUNCH_BATCHED) {
        // batch-global band numbering for the per-band path lists
        int64_t bands = 0;
        for (int32_t i = 0; i < b.n_pages; ++i) {
            if (b.pages[i].width > 32767)
                return fail(c, RR_E_INVARIANT, "batched launch: page width must be <= 32767");
            c->host_pages[static_cast<size_t>(i)].band_base = static_cast<int32_t>(bands);
            bands += (b.pages[i].height + kBandRows - 1) / kBandRows;
            if (bands > (1ll << 30)) return fail(c, RR_E_INVARIANT, "batched launch: too many bands");
        }
        // composite tile items: pages with more paths first, bands x tiles each
        const int32_t chunk_px = kVariantCH[b.variant];
        std::vector<int32_t> order(static_cast<size_t>(b.n_pages));
        for (int32_t i = 0; i < b.n_pages; ++i) order[static_cast<size_t>(i)] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return (b.pages[x].path_end - b.pages[x].path_begin) > (b.pages[y].path_end - b.pages[y].path_begin);
        });
        c->host_cpages.resize(static_cast<size_t>(b.n_pages));
        long long n_citems = 0;
        for (size_t k = 0; k < order.size(); ++k) {
            const rr_page_desc &pg = b.pages[order[k]];
            CPage &cp = c->host_cpages[k];
            cp.page = order[k];
            cp.tiles = (pg.width + chunk_px - 1) / chunk_px;
            cp.base = static_cast<int32_t>(n_citems);
            cp.pad = 0;
            DevPage &d = c->host_pages[static_cast<size_t>(order[k])];
            d.cbase = cp.base;
            d.ctiles = cp.tiles;
            n_citems += static_cast<long long>(cp.tiles) * ((pg.height + kBandRows - 1) / kBandRows);
            if (n_citems >= 0x7fffffffll) return fail(c, RR_E_INVARIANT, "batched launch: too many composite tiles");
        }
        c->batch_citems = n_citems;
    }
    if (c->capturing && b.launch != RR_LAUNCH_BATCHED)
        return fail(c, RR_E_INVARIANT, "graph capture covers the batched launch shape only");
    if (!c->capturing) {
        RR_HIP_TRY(c, hipMemcpyAsync(c->pages, c->host_pages.data(), sizeof(DevPage) * b.n_pages,
                                     hipMemcpyHostToDevice, c->stream));
        // The host tables are reused by the next recording: let the copies land.
        RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    }

    FillArgs A{};
    A.recs = c->recs;
    A.hdrs = c->hdrs;
    A.rows = c->rows;
    A.csum = c->csum;
    A.edge_off = b.edge_off;
    A.attrs = b.attrs;
    A.pages = c->pages;
    A.cov_hash = b.cov_hash;
    A.cov_count = b.cov_count;
    A.spill = c->spill;
    A.spill_top = c->spill_top;
    A.err = c->err_flag;
    A.spill_cap = c->spill_cap;
    A.mode = b.mode;
    A.flavour = b.flavour;
    A.colors = reinterpret_cast<const uint32_t *>(b.colors);
    A.slot_ints = kSlotInts;
    A.trace = c->trace;
    A.trace_cap = c->trace_cap;
    A.trace_top = c->trace_top;

    if (!c->page_timed && !c->capturing) {
        RR_HIP_TRY(c, hipEventRecord(c->t0, c->stream));
        c->page_timed = true;
    }
    rc = rrint::zero_u64(c, c->spill_top, 1);
    if (rc) return rc;
    // chunk summaries start empty: {INT32_MAX, INT32_MIN}
    if (n_chunks > 0) rc = rrint::fill_u64(c, reinterpret_cast<unsigned long long *>(c->csum), n_chunks,
                                            0x800000007fffffffull);
    if (rc) return rc;
    const unsigned setup_blocks =
        static_cast<unsigned>(std::min<long long>((b.n_paths + 3) / 4, static_cast<long long>(c->n_cu) * c->setup_bpc));
    hipLaunchKernelGGL(k_edge_setup, dim3(setup_blocks), dim3(256), 0, c->stream,
                       b.edges, b.edge_off, b.n_paths, c->pages, b.n_pages, c->recs, c->hdrs, c->rows, c->sel, c->nb,
                       b.attrs, b.launch == RR_LAUNCH_BATCHED ? c->feat : nullptr, c->csum, b.mode, c->err_flag);
    RR_HIP_TRY(c, hipGetLastError());

    if (b.launch == RR_LAUNCH_PLAIN) {
        const int NS = c->plain_streams;
        constexpr int CELL = rr_hip_ctx::kPlainCellRows;
        int32_t max_bands = 1;
        for (int32_t i = 0; i < b.n_pages; ++i)
            max_bands = std::max(max_bands, (b.pages[i].height + kBandRows) / kBandRows + 1);
        // one block of overflow slots per side stream (concurrent launches
        // never share a slot)
        rc = rrint::ensure(c, &c->slots, &c->slots_cap, static_cast<size_t>(NS) * max_bands * kSlotInts);
        if (rc) return rc;
        if (!c->side_start) {
            for (int s = 0; s < NS; ++s) {
                RR_HIP_TRY(c, hipStreamCreateWithFlags(&c->side[s], hipStreamNonBlocking));
                RR_HIP_TRY(c, hipEventCreateWithFlags(&c->side_ev[s], hipEventDisableTiming));
            }
            RR_HIP_TRY(c, hipEventCreateWithFlags(&c->side_start, hipEventDisableTiming));
        }
        A.slots = c->slots;
        if (!c->plain_args) RR_HIP_TRY(c, hipMalloc(&c->plain_args, sizeof(FillArgs)));
        // ordered after the launches of earlier recordings (the context stream
        // joined their side streams) and ahead of this one's
        hipLaunchKernelGGL(k_store_fill_args, dim3(1), dim3(64), 0, c->stream, A, c->plain_args);
        RR_HIP_TRY(c, hipGetLastError());
        // the side streams start after the work recorded on the context stream
        RR_HIP_TRY(c, hipEventRecord(c->side_start, c->stream));
        uint32_t started = 0;
        auto start = [&](int s) -> int32_t {
            if (!((started >> s) & 1u)) {
                RR_HIP_TRY(c, hipStreamWaitEvent(c->side[s], c->side_start, 0));
                started |= 1u << s;
            }
            return RR_OK;
        };
        auto join_all = [&](hipStream_t into) -> int32_t {
            for (int s = 0; s < NS; ++s) {
                if (!((started >> s) & 1u)) continue;
                RR_HIP_TRY(c, hipEventRecord(c->side_ev[s], c->side[s]));
                RR_HIP_TRY(c, hipStreamWaitEvent(into, c->side_ev[s], 0));
            }
            return RR_OK;
        };
        int next = 0;
        for (int32_t pi = 0; pi < b.n_pages; ++pi) {
            const rr_page_desc &pg = b.pages[pi];
            c->cell_stream.assign(static_cast<size_t>((pg.height + CELL - 1) / CELL), static_cast<int8_t>(-1));
            for (int64_t p = pg.path_begin; p < pg.path_end; ++p) {
                int32_t r0 = b.host_rows[2 * p], r1 = b.host_rows[2 * p + 1];
                r0 = std::max(r0, 0);
                r1 = std::min(r1, pg.height - 1);
                if (r0 > r1) continue;
                // side streams that drew a row cell of this path last
                uint32_t dep = 0;
                for (int32_t q = r0 / CELL; q <= r1 / CELL; ++q)
                    if (c->cell_stream[q] >= 0) dep |= 1u << c->cell_stream[q];
                int s;
                if (dep == 0) {
                    s = next;
                    // per-launch profiling runs the launches one after
                    // another on the first stream, so that each event pair
                    // brackets its kernel alone
                    next = c->profiling ? 0 : (next + 1) % NS;
                } else {
                    s = __builtin_ctz(dep);
                }
                rc = start(s);
                if (rc) return rc;
                for (uint32_t o = dep & ~(1u << s); o; o &= o - 1) {
                    const int t = __builtin_ctz(o);
                    RR_HIP_TRY(c, hipEventRecord(c->side_ev[t], c->side[t]));
                    RR_HIP_TRY(c, hipStreamWaitEvent(c->side[s], c->side_ev[t], 0));
                }
                const unsigned bands = static_cast<unsigned>((r1 - r0) / kBandRows + 1);
                const int32_t slot0 = s * max_bands;
                // two or three bands: one workgroup (the device dispatches
                // single-workgroup launches at a higher rate)
                const int shape = bands == 2 ? 1 : (bands == 3 ? 2 : 0);
                const dim3 grid(shape ? 1u : bands), block(shape ? 64u * kPlainGroupWaves[shape] * bands : 64u * kMwWaves);
                hipEvent_t *ev = c->profiling ? prof_pair(c) : nullptr;
                if (c->profiling && !ev) return fail(c, RR_E_DEVICE, "profiling event allocation failed");
                if (ev) RR_HIP_TRY(c, hipEventRecord(ev[0], c->side[s]));
                const FillArgs *pa = c->plain_args;
                void *args[] = {&pa, const_cast<int32_t *>(&pi), &p, &r0, const_cast<int32_t *>(&slot0)};
                RR_HIP_TRY(c, hipLaunchKernel(kPlainKernels[shape][b.variant], grid, block, args, 0, c->side[s]));
                if (ev) RR_HIP_TRY(c, hipEventRecord(ev[1], c->side[s]));
                ++c->fill_launches;
                for (int32_t q = r0 / CELL; q <= r1 / CELL; ++q) c->cell_stream[q] = static_cast<int8_t>(s);
            }
            if (pg.gray) {
                // the page's Gray8 output once its last path is composited
                rc = join_all(c->stream);
                if (!rc) rc = rrint::rgb_to_gray(c, pg.canvas, pg.stride, pg.gray, pg.gray_stride, pg.width, pg.height);
                if (rc) return rc;
                RR_HIP_TRY(c, hipEventRecord(c->side_start, c->stream));
                started = 0;
            }
        }
        // the context stream (the page fence, later recordings) waits for
        // every side stream
        rc = join_all(c->stream);
        if (rc) return rc;
        RR_HIP_TRY(c, hipGetLastError());
        return RR_OK;
    }

    return rrint::record_batched(c, b, A, kVariantK[b.variant], kVariantCH[b.variant], b.variant);
}

}  // namespace

namespace rrint {
int32_t record_pages_internal(rr_hip_ctx *c, const rr_fill_pages &b) { return record_pages(c, b); }
hipStream_t stream(rr_hip_ctx *c) { return c->stream; }
int device(rr_hip_ctx *c) { return c->device; }
bool page_open(rr_hip_ctx *c) { return c->page_open; }
int32_t fail(rr_hip_ctx *c, int32_t code, const std::string &msg) { return ::fail(c, code, msg); }
hipEvent_t *prof_pair(rr_hip_ctx *c, bool *ok) {
    *ok = true;
    if (!c->profiling) return nullptr;
    hipEvent_t *e = ::prof_pair(c);
    *ok = e != nullptr;
    return e;
}
void count_launch(rr_hip_ctx *c) { ++c->fill_launches; }
}  // namespace rrint

extern "C" {

int32_t rr_hip_create(int32_t device, rr_hip_ctx **out) {
    if (!out) return RR_E_INVARIANT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return RR_E_NO_DEVICE;
    rr_hip_ctx *c = new rr_hip_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RR_E_NO_DEVICE;
    }
    int32_t rc = RR_OK;
    if (hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) rc = RR_E_DEVICE;
    if (rc == RR_OK) rc = rrint::init_batched(c);
    if (rc == RR_OK &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&c->setup_bpc, reinterpret_cast<const void *>(&k_edge_setup), 256,
                                                     0) != hipSuccess)
        rc = RR_E_DEVICE;
    c->setup_bpc = std::max(1, c->setup_bpc);
    // side streams of the plain launch: one hardware queue each, two left for
    // the context stream and HIP's own (HIP puts streams beyond its queue
    // count on shared queues, which serialises them); with few queues (HIP's
    // default is 4) every queue is shared anyway and the full set keeps the
    // most launches in flight
    {
        const char *q = std::getenv("GPU_MAX_HW_QUEUES");
        const int hwq = q ? std::atoi(q) : 4;
        c->plain_streams = hwq >= 8 ? std::min(rr_hip_ctx::kPlainStreams, hwq - 2) : rr_hip_ctx::kPlainStreams;
    }
    if (rc == RR_OK) rc = rrint::init_static_scratch(c);
    if (rc == RR_OK && (hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess)) rc = RR_E_DEVICE;
    if (rc == RR_OK && hipMalloc(&c->spill_top, sizeof(unsigned long long)) != hipSuccess) rc = RR_E_DEVICE;
    if (rc == RR_OK && hipMalloc(&c->counters, kCounters * sizeof(unsigned long long)) != hipSuccess) rc = RR_E_DEVICE;
    if (rc == RR_OK && hipMalloc(&c->queues, kQueueWords * sizeof(unsigned long long)) != hipSuccess) rc = RR_E_DEVICE;
    if (rc == RR_OK && hipMalloc(&c->err_flag, (1 + kFenceSlots) * sizeof(int32_t)) != hipSuccess) rc = RR_E_DEVICE;
    if (rc == RR_OK && hipMalloc(&c->trace_top, sizeof(unsigned long long)) != hipSuccess) rc = RR_E_DEVICE;
    for (auto &f : c->fences)
        if (rc == RR_OK && hipEventCreateWithFlags(&f.ev, hipEventDisableTiming) != hipSuccess) rc = RR_E_DEVICE;
    const long long pool_ints = 64ll << 20;  // 256 MiB shared pool for rows beyond a wave's slot
    if (rc == RR_OK && hipMalloc(&c->spill, pool_ints * sizeof(int32_t)) != hipSuccess) rc = RR_E_DEVICE;
    c->spill_cap = pool_ints;
    if (rc == RR_OK && hipMemset(c->err_flag, 0, (1 + kFenceSlots) * sizeof(int32_t)) != hipSuccess) rc = RR_E_DEVICE;
    if (rc != RR_OK) {
        rr_hip_destroy(c);
        return rc;
    }
    *out = c;
    return RR_OK;
}

int32_t rr_hip_destroy(rr_hip_ctx *c) {
    if (!c) return RR_E_INVARIANT;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < rr_hip_ctx::kPlainStreams; ++k)
        if (c->side[k]) (void)hipStreamSynchronize(c->side[k]);
    if (c->aux) (void)hipStreamSynchronize(c->aux);
    void *bufs[] = {c->recs,      c->hdrs,     c->rows,       c->sel,     c->nb,        c->csum,
                    c->pages,     c->slots,    c->spill,      c->spill_top, c->err_flag, c->trace_top,
                    c->item_base, c->scan_tmp, c->sitems,     c->shdr,    c->spans,     c->band_cnt,
                    c->band_base, c->band_list, c->cpages,    c->counters, c->occ,      c->pstate,
                    c->feat,      c->route_cls, c->n_work,    c->work_base, c->mlp,       c->spath,
                    c->ovq,       c->bin_cnt,  c->bin_off,    c->bins,    c->queues,    c->band_run,
                    c->xp_t,      c->xp_role,  c->xp_close,   c->xp_cnt,  c->xp_eoff,   c->xp_err,
                    c->xp_sidx,   c->xp_seg,   c->img_grey,   c->xp_stage,  c->xp_bkt,
                    c->xp_perm,   c->xp_ccnt,  c->xp_coff,   c->plain_args};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto e : c->xp_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    for (auto &f : c->fences)
        if (f.ev) (void)hipEventDestroy(f.ev);
    for (auto e : c->prof_ev) (void)hipEventDestroy(e);
    for (int k = 0; k < rr_hip_ctx::kPlainStreams; ++k) {
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
        if (c->side_ev[k]) (void)hipEventDestroy(c->side_ev[k]);
    }
    if (c->side_start) (void)hipEventDestroy(c->side_start);
    if (c->aux) (void)hipStreamDestroy(c->aux);
    if (c->aux_fork) (void)hipEventDestroy(c->aux_fork);
    if (c->aux_join) (void)hipEventDestroy(c->aux_join);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RR_OK;
}

const char *rr_hip_last_error(rr_hip_ctx *c) { return c ? c->err.c_str() : "null context"; }

int32_t rr_hip_alloc(rr_hip_ctx *c, size_t size, int32_t zeroed, void **buf) {
    if (!c || !buf) return RR_E_INVARIANT;
    if (size == 0) return fail(c, RR_E_ZERO_SIZE_ALLOC, "rr_hip_alloc called with size = 0");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    RR_HIP_TRY(c, hipMalloc(buf, size));
    if (zeroed) RR_HIP_TRY(c, hipMemsetAsync(*buf, 0, size, c->stream));
    c->buf_len[*buf] = size;
    return RR_OK;
}

int32_t rr_hip_free(rr_hip_ctx *c, void *buf) {
    if (!c) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    RR_HIP_TRY(c, hipFree(buf));
    c->buf_len.erase(buf);
    return RR_OK;
}

int32_t rr_hip_buffer_len(rr_hip_ctx *c, const void *buf, size_t *len) {
    if (!c || !len) return RR_E_INVARIANT;
    const auto it = c->buf_len.find(buf);
    if (it == c->buf_len.end()) return fail(c, RR_E_INVARIANT, "buffer_len: not a buffer from rr_hip_alloc");
    *len = it->second;
    return RR_OK;
}

int32_t rr_hip_vram_budget(rr_hip_ctx *c, uint64_t *total_bytes, uint64_t *usable_bytes) {
    if (!c || !total_bytes || !usable_bytes) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipSetDevice(c->device));
    size_t free_b = 0, total_b = 0;
    RR_HIP_TRY(c, hipMemGetInfo(&free_b, &total_b));
    *total_bytes = total_b;
    *usable_bytes = static_cast<uint64_t>(free_b) / 4 * 3 + static_cast<uint64_t>(free_b) % 4 * 3 / 4;
    return RR_OK;
}

int32_t rr_hip_alloc_host_pinned(rr_hip_ctx *c, size_t size, void **buf) {
    if (!c || !buf) return RR_E_INVARIANT;
    *buf = nullptr;
    if (size == 0) return fail(c, RR_E_ZERO_SIZE_ALLOC, "alloc_host_pinned called with size = 0");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    RR_HIP_TRY(c, hipHostMalloc(buf, size, hipHostMallocDefault));
    return RR_OK;
}

int32_t rr_hip_free_host_pinned(rr_hip_ctx *c, void *buf) {
    if (!c) return RR_E_INVARIANT;
    if (!buf) return RR_OK;
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    RR_HIP_TRY(c, hipHostFree(buf));
    return RR_OK;
}

int32_t rr_hip_upload(rr_hip_ctx *c, void *dst, const void *src, size_t n) {
    if (!c || (n && (!dst || !src))) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RR_OK;
}

int32_t rr_hip_download(rr_hip_ctx *c, void *dst, const void *src, size_t n) {
    if (!c || (n && (!dst || !src))) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RR_OK;
}

int32_t rr_hip_copy(rr_hip_ctx *c, void *dst, const void *src, size_t n) {
    if (!c || (n && (!dst || !src))) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, c->stream));
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    return RR_OK;
}

int32_t rr_hip_memset(rr_hip_ctx *c, void *dst, int32_t value, size_t n) {
    if (!c || (n && !dst)) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipMemsetAsync(dst, value, n, c->stream));
    return RR_OK;
}

int32_t rr_hip_begin_page(rr_hip_ctx *c) {
    if (!c) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipSetDevice(c->device));
    const int32_t rc = rrint::fill_i32(c, c->err_flag, 1, 0);  // the live word; fence slots keep theirs
    if (rc) return rc;
    c->page_open = true;
    c->page_timed = false;
    c->fill_launches = 0;
    c->prof_used = 0;
    return RR_OK;
}

int32_t rr_hip_inject_error(rr_hip_ctx *c, int32_t bits) {
    if (!c || bits == 0) return RR_E_INVARIANT;
    if (!c->page_open) return fail(c, RR_E_INVARIANT, "inject_error outside begin_page/submit_page");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    return rrint::set_err_bits(c, bits);
}

int32_t rr_hip_record_fill_pages(rr_hip_ctx *c, const rr_fill_pages *b) {
    if (!c || !b) return RR_E_INVARIANT;
    return record_pages(c, *b);
}

int32_t rr_hip_record_fill_paths(rr_hip_ctx *c, const rr_fill_batch *b) {
    if (!c || !b) return RR_E_INVARIANT;
    rr_page_desc pg{};
    pg.canvas = b->canvas;
    pg.width = b->width;
    pg.height = b->height;
    pg.stride = b->stride;
    pg.path_begin = 0;
    pg.path_end = b->n_paths;
    for (int k = 0; k < 4; ++k) pg.clip[k] = b->clip[k];
    rr_fill_pages fp{};
    fp.edges = b->edges;
    fp.edge_off = b->edge_off;
    fp.attrs = b->attrs;
    fp.host_rows = b->host_rows;
    fp.n_paths = b->n_paths;
    fp.n_edges = b->n_edges;
    fp.pages = &pg;
    fp.n_pages = 1;
    fp.mode = b->mode;
    fp.flavour = b->flavour;
    fp.launch = b->launch;
    fp.cov_hash = b->cov_hash;
    fp.cov_count = b->cov_count;
    fp.variant = b->variant;
    fp.colors = b->colors;
    pg.soft_mask = b->soft_mask;
    pg.soft_mask_stride = b->soft_mask_stride;
    pg.gray = b->gray;
    pg.gray_stride = b->gray_stride;
    return record_pages(c, fp);
}

int32_t rr_hip_submit_page(rr_hip_ctx *c, uint64_t *fence) {
    if (!c || !fence) return RR_E_INVARIANT;
    if (!c->page_open) return fail(c, RR_E_INVARIANT, "submit_page without begin_page");
    if (c->page_timed) RR_HIP_TRY(c, hipEventRecord(c->t1, c->stream));
    const int slot = static_cast<int>(c->next_fence % kFenceSlots);
    Fence &f = c->fences[slot];
    // the page's error word goes with its fence: the next begin_page clears
    // only the live word
    RR_HIP_TRY(c, hipMemcpyAsync(c->err_flag + 1 + slot, c->err_flag, sizeof(int32_t), hipMemcpyDeviceToDevice,
                                 c->stream));
    RR_HIP_TRY(c, hipEventRecord(f.ev, c->stream));
    f.id = c->next_fence;
    f.page = true;
    *fence = c->next_fence++;
    c->page_open = false;
    c->last_launches = c->fill_launches;
    return RR_OK;
}

// Blocks until fence f (a page or transfer fence of this context) has
// signalled.  *slot receives its ring slot, or -1 for a fence older than the
// ring (complete, but its page's error word is no longer held).
static int32_t wait_fence(rr_hip_ctx *c, uint64_t fence, int *slot) {
    if (fence == 0 || fence >= c->next_fence) return fail(c, RR_E_INVARIANT, "wait: unknown fence");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    Fence &f = c->fences[fence % kFenceSlots];
    if (f.id != fence) {
        RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
        *slot = -1;
    } else {
        RR_HIP_TRY(c, hipEventSynchronize(f.ev));
        *slot = static_cast<int>(fence % kFenceSlots);
    }
    return RR_OK;
}

int32_t rr_hip_wait_page(rr_hip_ctx *c, uint64_t fence) {
    if (!c) return RR_E_INVARIANT;
    int slot = -1;
    const int32_t wrc = wait_fence(c, fence, &slot);
    if (wrc) return wrc;
    if (slot < 0) return fail(c, RR_E_INVARIANT, "wait_page: fence older than the ring; its page status is gone");
    if (!c->fences[slot].page) return fail(c, RR_E_INVARIANT, "wait_page: a transfer fence");
    int32_t flag = 0;
    RR_HIP_TRY(c, hipMemcpy(&flag, c->err_flag + 1 + slot, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (c->page_timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->t0, c->t1) == hipSuccess) c->last_ms = ms;
    }
    double fill = 0.0;
    for (size_t k = 0; k + 1 < c->prof_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->prof_ev[k], c->prof_ev[k + 1]) == hipSuccess) fill += ms;
    }
    c->last_fill_ms = fill;
    if (flag & kErrNoSoftMask)
        return fail(c, RR_E_INVARIANT, "a path asked for the soft mask of a page that has none (filled unmasked)");
    if (flag & 1) return fail(c, RR_E_CAPACITY, "overflow-entry pool exhausted: page output is incomplete");
    if (flag) return fail(c, RR_E_CAPACITY, "batched workspace overflow (code " + std::to_string(flag) + "): page output is incomplete");
    return RR_OK;
}

// Transfer fence: an event on the context stream after everything recorded
// so far (uploads, downloads and pages are ordered on that one stream).
static int32_t transfer_fence(rr_hip_ctx *c, uint64_t *fence) {
    Fence &f = c->fences[c->next_fence % kFenceSlots];
    RR_HIP_TRY(c, hipEventRecord(f.ev, c->stream));
    f.id = c->next_fence;
    f.page = false;
    *fence = c->next_fence++;
    return RR_OK;
}

int32_t rr_hip_upload_async(rr_hip_ctx *c, void *dst, const void *src, size_t n, uint64_t *fence) {
    if (!c || !fence || (n && (!dst || !src))) return RR_E_INVARIANT;
    const auto it = c->buf_len.find(dst);
    if (it != c->buf_len.end() && n > it->second) return fail(c, RR_E_INVARIANT, "upload_async: source longer than buffer");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    if (n) RR_HIP_TRY(c, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream));
    return transfer_fence(c, fence);
}

int32_t rr_hip_download_async(rr_hip_ctx *c, void *dst, const void *src, size_t n, uint64_t *fence) {
    if (!c || !fence || (n && (!dst || !src))) return RR_E_INVARIANT;
    const auto it = c->buf_len.find(src);
    if (it != c->buf_len.end() && n > it->second) return fail(c, RR_E_INVARIANT, "download_async: longer than buffer");
    RR_HIP_TRY(c, hipSetDevice(c->device));
    if (n) RR_HIP_TRY(c, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, c->stream));
    return transfer_fence(c, fence);
}

int32_t rr_hip_submit_transfer(rr_hip_ctx *c, uint64_t *fence) {
    if (!c || !fence) return RR_E_INVARIANT;
    RR_HIP_TRY(c, hipSetDevice(c->device));
    return transfer_fence(c, fence);
}

int32_t rr_hip_wait_transfer(rr_hip_ctx *c, uint64_t fence) {
    if (!c) return RR_E_INVARIANT;
    int slot = -1;
    return wait_fence(c, fence, &slot);
}

int32_t rr_hip_variant_info(int32_t variant, int32_t *lds_entries, int32_t *chunk_px) {
    if (variant < 0 || variant >= kNumVariants) return RR_E_INVARIANT;
    if (lds_entries) *lds_entries = kVariantK[variant];
    if (chunk_px) *chunk_px = kVariantCH[variant];
    return RR_OK;
}

int32_t rr_hip_variant_count(void) { return kNumVariants; }

int32_t rr_hip_autotune_fill(rr_hip_ctx *c, const rr_fill_pages *b, int32_t reps, int32_t *best, double *ms) {
    if (!c || !b || !best || reps <= 0) return RR_E_INVARIANT;
    if (c->page_open) return fail(c, RR_E_INVARIANT, "autotune inside an open page");
    rr_fill_pages t = *b;
    t.cov_hash = nullptr;
    t.cov_count = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RR_HIP_TRY(c, hipEventCreate(&e0));
    RR_HIP_TRY(c, hipEventCreate(&e1));
    double best_ms = 0.0;
    int32_t rc = RR_OK;
    *best = 0;
    for (int v = 0; v < kNumVariants && rc == RR_OK; ++v) {
        t.variant = v;
        double sum = 0.0;
        for (int r = 0; r <= reps && rc == RR_OK; ++r) {  // r == 0 is a warm-up
            rr_hip_begin_page(c);
            if (hipEventRecord(e0, c->stream) != hipSuccess) rc = RR_E_DEVICE;
            if (rc == RR_OK) rc = record_pages(c, t);
            if (rc == RR_OK && hipEventRecord(e1, c->stream) != hipSuccess) rc = RR_E_DEVICE;
            uint64_t f = 0;
            const int32_t rs = rr_hip_submit_page(c, &f);
            if (rc == RR_OK) rc = rs;
            if (rc == RR_OK) rc = rr_hip_wait_page(c, f);
            float m = 0.f;
            if (rc == RR_OK && hipEventElapsedTime(&m, e0, e1) != hipSuccess) rc = RR_E_DEVICE;
            if (r > 0) sum += m;
        }
        const double avg = sum / reps;
        if (ms) ms[v] = avg;
        if (rc == RR_OK && (v == 0 || avg < best_ms)) {
            best_ms = avg;
            *best = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int32_t rr_hip_set_trace(rr_hip_ctx *c, void *records, int64_t capacity) {
    if (!c || capacity < 0 || (capacity > 0 && !records)) return RR_E_INVARIANT;
    if (c->page_open) return fail(c, RR_E_INVARIANT, "set_trace inside an open page");
    if (reinterpret_cast<uintptr_t>(records) & 31u) return fail(c, RR_E_INVARIANT, "trace buffer must be 32-byte aligned");
    c->trace = capacity ? static_cast<longlong4 *>(records) : nullptr;
    c->trace_cap = capacity;
    RR_HIP_TRY(c, hipMemsetAsync(c->trace_top, 0, sizeof(unsigned long long), c->stream));
    return RR_OK;
}

int32_t rr_hip_trace_count(rr_hip_ctx *c, int64_t *count) {
    if (!c || !count) return RR_E_INVARIANT;
    unsigned long long n = 0;
    RR_HIP_TRY(c, hipStreamSynchronize(c->stream));
    RR_HIP_TRY(c, hipMemcpy(&n, c->trace_top, sizeof(n), hipMemcpyDeviceToHost));
    *count = static_cast<int64_t>(n);
    return RR_OK;
}

int32_t rr_hip_set_profiling(rr_hip_ctx *c, int32_t enable) {
    if (!c) return RR_E_INVARIANT;
    if (c->page_open) return fail(c, RR_E_INVARIANT, "set_profiling inside an open page");
    c->profiling = enable != 0;
    return RR_OK;
}

int32_t rr_hip_last_page_stats(rr_hip_ctx *c, double *span_ms, double *fill_ms, int64_t *launches) {
    if (!c) return RR_E_INVARIANT;
    if (span_ms) *span_ms = c->last_ms;
    if (fill_ms) *fill_ms = c->last_fill_ms;
    if (launches) *launches = c->last_launches;
    return RR_OK;
}

}  // extern "C"
