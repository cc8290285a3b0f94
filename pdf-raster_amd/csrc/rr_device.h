// rr_device.h -- device-side data layout and scanline helpers shared by the
// fill kernels (plain and batched routes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rr_hip.h"

namespace rrdev {

// ---------------------------------------------------------------------------
// device data layout
// ---------------------------------------------------------------------------

// One edge after add_segment + aa_scale, reduced to what the per-row integer
// stepping needs.  32 bytes, so one wave loads 64 of them in 2 KB.
struct EdgeRec {
    int32_t r0, r1;  // AA rows [r0, r1] after the clip; empty when r0 > r1
    int32_t meta;    // bits 0-1 kind (0 sloped, 1 vertical, 2 horizontal),
                     // bit 2 count +1 (FLIPPED), bit 3 first row has count 0,
                     // bit 4 EDGE_FAST (sloped, no saturation over its rows)
    int32_t a;       // sloped: px0 of row r0; vertical: x; horizontal: x0
    int32_t b;       // sloped: 16.16 right edge at row r0; horizontal: x1
    int32_t c;       // sloped: dxdy_fp
    int32_t d, e;    // sloped: 16.16 clamp min / max
};
static_assert(sizeof(EdgeRec) == 32, "EdgeRec layout");

enum { KIND_SLOPED = 0, KIND_VERT = 1, KIND_HORIZ = 2 };
constexpr int32_t EDGE_FAST = 16;
enum { HDR_EMPTY = 1, HDR_OUTSIDE = 2 };

// Scanner bounding box of one path in AA space (scanner/mod.rs:138-169).
struct PathHdr {
    int32_t y_min, y_max, x_min, x_max;
    int32_t flags, page;
    int32_t ref_area;  // the reference dispatcher's bbox area in pixels, -1 when it keeps the fill on the CPU
    int32_t pad2;
};

struct ClipRect {
    double x_min, y_min, x_max, y_max;
    int32_t x_min_i, y_min_i, x_max_i, y_max_i;
};

// One page of a batch, as the kernels see it.
struct DevPage {
    uint8_t *canvas;
    int64_t path_begin, path_end;
    ClipRect clip;
    int32_t width, height, stride;
    int32_t px_lo, px_hi;  // clip ∩ canvas, pixels
    int32_t band_base;     // batch-global index of the page's first 16-row band
    int32_t cbase, ctiles;  // batched: first composite tile item of the page, tiles per band
    int32_t pad2;
    const uint8_t *smask;  // page-space soft mask (RR_FILL_SOFT_MASK paths) or null
    long long smask_stride;
    uint8_t *gray;  // RR_FLAVOUR_RGB8: Gray8 output of the finished page, or null
    long long gray_stride;
};

struct FillArgs {
    const EdgeRec *recs;
    const PathHdr *hdrs;
    const int2 *rows;  // per path: drawn output rows [lo, hi] (lo > hi: nothing)
    const int2 *csum;  // per 64-edge chunk (global edge index / 64): min r0, max r1
    const int64_t *edge_off;
    const rr_path_attr *attrs;
    const DevPage *pages;
    uint64_t *cov_hash;
    uint64_t *cov_count;
    int32_t *slots;                 // per-wave overflow entry slots
    long long slot_ints;            // ints per slot
    int32_t *spill;                 // shared overflow pool for rows beyond a slot
    unsigned long long *spill_top;  // bump allocator (in int32 units)
    int32_t *err;                   // set to 1 when the pool runs out
    long long spill_cap;
    int32_t mode, flavour;
    const uint32_t *colors;  // RR_FLAVOUR_RGB8: per path r | g << 8 | b << 16
    // optional timing trace: one record per (path, band) fill
    longlong4 *trace;  // {path, (page << 32) | first row, t_start, t_end} wall-clock ticks
    unsigned long long *trace_top;
    long long trace_cap;
};

// AA_GAMMA (crates/raster/src/fill/mod.rs:46-48)
// ---------------------------------------------------------------------------
// wave64 cross-lane helpers on DPP (no LDS round trip).  All lanes of the
// wave must be active where they are used.
// ---------------------------------------------------------------------------
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;
constexpr int kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E;  // quad_perm [1,0,3,2], [2,3,0,1]

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int32_t dpp_i32(int32_t v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, true);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_update_dpp(0, static_cast<int32_t>(v), CTRL, ROW_MASK, 0xf, true));
}

// DPP move whose lanes without a source (or outside ROW_MASK) keep `keep`.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int32_t dpp_keep(int32_t v, int32_t keep) {
    return __builtin_amdgcn_update_dpp(keep, v, CTRL, ROW_MASK, 0xf, false);
}

// Value of lane (lane & ~3) + K, the K-th lane of this lane's quad.
template <int K>
__device__ __forceinline__ int32_t quad_bcast(int32_t v) {
    return dpp_i32<0x55 * K>(v);
}

// Lane 0's value of a 64-bit quantity, in scalar registers.
__device__ __forceinline__ unsigned long long first_lane_u64(unsigned long long v) {
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(v)));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(v >> 32)));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Inclusive prefix sum over the wave.
__device__ __forceinline__ int32_t wave_incl_add(int32_t v) {
    v += dpp_i32<kDppRowShr1>(v);
    v += dpp_i32<kDppRowShr2>(v);
    v += dpp_i32<kDppRowShr4>(v);
    v += dpp_i32<kDppRowShr8>(v);
    v += dpp_i32<kDppRowBcast15, 0xa>(v);
    v += dpp_i32<kDppRowBcast31, 0xc>(v);
    return v;
}

// Maximum / minimum over the wave (a DPP scan to lane 63), wave-uniform.
__device__ __forceinline__ int32_t wave_max(int32_t v) {
    v = max(v, dpp_keep<kDppRowShr1>(v, INT32_MIN));
    v = max(v, dpp_keep<kDppRowShr2>(v, INT32_MIN));
    v = max(v, dpp_keep<kDppRowShr4>(v, INT32_MIN));
    v = max(v, dpp_keep<kDppRowShr8>(v, INT32_MIN));
    v = max(v, dpp_keep<kDppRowBcast15, 0xa>(v, INT32_MIN));
    v = max(v, dpp_keep<kDppRowBcast31, 0xc>(v, INT32_MIN));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int32_t wave_min(int32_t v) {
    v = min(v, dpp_keep<kDppRowShr1>(v, INT32_MAX));
    v = min(v, dpp_keep<kDppRowShr2>(v, INT32_MAX));
    v = min(v, dpp_keep<kDppRowShr4>(v, INT32_MAX));
    v = min(v, dpp_keep<kDppRowShr8>(v, INT32_MAX));
    v = min(v, dpp_keep<kDppRowBcast15, 0xa>(v, INT32_MAX));
    v = min(v, dpp_keep<kDppRowBcast31, 0xc>(v, INT32_MAX));
    return __builtin_amdgcn_readlane(v, 63);
}

// Inclusive prefix OR over the wave.
__device__ __forceinline__ uint32_t wave_incl_or(uint32_t v) {
    v |= dpp_u32<kDppRowShr1>(v);
    v |= dpp_u32<kDppRowShr2>(v);
    v |= dpp_u32<kDppRowShr4>(v);
    v |= dpp_u32<kDppRowShr8>(v);
    v |= dpp_u32<kDppRowBcast15, 0xa>(v);
    v |= dpp_u32<kDppRowBcast31, 0xc>(v);
    return v;
}

// Sum over the wave of an f64 by an xor butterfly: every lane ends with the
// same value (each step adds the same two operands, in either order).
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix XOR over the wave, 64-bit.
__device__ __forceinline__ uint32_t wave_incl_xor(uint32_t v) {
    v ^= dpp_u32<kDppRowShr1>(v);
    v ^= dpp_u32<kDppRowShr2>(v);
    v ^= dpp_u32<kDppRowShr4>(v);
    v ^= dpp_u32<kDppRowShr8>(v);
    v ^= dpp_u32<kDppRowBcast15, 0xa>(v);
    v ^= dpp_u32<kDppRowBcast31, 0xc>(v);
    return v;
}

__device__ __forceinline__ unsigned long long wave_incl_xor64(unsigned long long v) {
    const uint32_t lo = wave_incl_xor(static_cast<uint32_t>(v)), hi = wave_incl_xor(static_cast<uint32_t>(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

__device__ __forceinline__ unsigned long long wave_incl_or64(unsigned long long v) {
    const uint32_t lo = wave_incl_or(static_cast<uint32_t>(v)), hi = wave_incl_or(static_cast<uint32_t>(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Value of lane - 1 (0 for lane 0).
__device__ __forceinline__ unsigned long long wave_shr1_u64(unsigned long long v) {
    const uint32_t lo = dpp_u32<kDppWaveShr1>(static_cast<uint32_t>(v));
    const uint32_t hi = dpp_u32<kDppWaveShr1>(static_cast<uint32_t>(v >> 32));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Error-word bits a kernel sets when a page's output is incomplete (1, 2, 4,
// 8, 16, 32, 64: workspace capacity) or its input is inconsistent (128: a
// path asks for the soft mask of a page without one).
constexpr int32_t kErrNoSoftMask = 128;

static __constant__ uint8_t kGamma[17] = {0, 4, 11, 21, 32, 45, 59, 74, 90, 108, 126, 145, 166, 187, 209, 231, 255};

// ---------------------------------------------------------------------------
// exact f64 helpers (crates/color/src/convert.rs:221-264; scanner/mod.rs:51-66)
// ---------------------------------------------------------------------------
// v_cvt_i32_f64: truncation toward zero, values beyond the int32 range (the
// infinities included) clamp to INT32_MIN / INT32_MAX, NaN gives 0 -- the
// saturating f64 -> i32 conversion of the reference's `as i32`, in one
// instruction.
__device__ __forceinline__ int32_t cvt_i32_sat(double x) {
    int32_t r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// floor, saturated to int32; NaN gives INT32_MIN
__device__ __forceinline__ int32_t splash_floor(double x) {
    const int32_t f = cvt_i32_sat(floor(x));
    return isnan(x) ? INT32_MIN : f;
}

// 16.16 fixed point, truncated toward zero and saturated; NaN gives 0
__device__ __forceinline__ int32_t fp_from_f64(double x) { return cvt_i32_sat(x * 65536.0); }

__device__ __forceinline__ uint64_t bits_of(double v) { return __double_as_longlong(v); }

__device__ __forceinline__ int32_t sat64_to_i32(long long v) {
    return v > INT32_MAX ? INT32_MAX : (v < INT32_MIN ? INT32_MIN : static_cast<int32_t>(v));
}

__device__ __forceinline__ uint32_t div255(uint32_t x) {
    const uint32_t s = (x + (x >> 8) + 0x80u) >> 8;
    return s < 255u ? s : 255u;
}

// The general pipe's share of a path's fill flags (rr_path_attr.eo bits):
// with either, the paint is compute_a_src + the no-alpha general composite.
constexpr uint32_t kFillGeneral = RR_FILL_MULTIPLY | RR_FILL_SOFT_MASK;

// One channel of the fill's composite (pipe/mod.rs:157-199): the general
// pipe's no-alpha arm (pipe/general.rs:270-344: a_src = compute_a_src with a
// shape; Normal c = div255((255 - a) d + a s), Multiply with blend_multiply
// = div255(s d), blend.rs:24-26) when the path is Multiply or soft-masked,
// else the AA path: Gray8 exact div255 (pipe/aa.rs:192-208) or the Rgb8 fast
// (v + 255) >> 8 (simd/composite.rs:56-114).
// Products of bytes go through v_mul_u32_u24: with both factors known to be
// bytes the compiler may otherwise fold div255(s * d) into a dot-product
// sequence whose results differ from the formula's.
__device__ __forceinline__ uint32_t mul24_u(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint32_t paint_px(uint32_t dv, uint32_t shape, uint32_t a_in, uint32_t src, uint32_t fl,
                                             uint32_t sm, bool fast) {
    if (fl & kFillGeneral) {
        const uint32_t a = (fl & RR_FILL_SOFT_MASK) ? div255(mul24_u(div255(mul24_u(a_in, sm)), shape))
                                                    : div255(mul24_u(a_in, shape));
        const uint32_t cb = (fl & RR_FILL_MULTIPLY) ? div255(mul24_u(src, dv)) : src;
        return div255(mul24_u(255u - a, dv) + mul24_u(a, cb));
    }
    if (fast) {
        const uint32_t a = (a_in * shape + 255u) >> 8;
        return ((255u - a) * dv + a * src + 255u) >> 8;
    }
    const uint32_t a = div255(a_in * shape);
    return div255((255u - a) * dv + a * src);
}

// rgb_to_gray (crates/pdf_raster/src/render.rs:1306-1324) of 4 pixels held
// as three channel dwords.
__device__ __forceinline__ uint32_t gray4(uint32_t r, uint32_t g, uint32_t b) {
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t v = (2126u * ((r >> (8 * k)) & 0xffu) + 7152u * ((g >> (8 * k)) & 0xffu) +
                            722u * ((b >> (8 * k)) & 0xffu) + 5000u) / 10000u;
        out |= v << (8 * k);
    }
    return out;
}

// 4 interleaved Rgb8 pixels (three dwords w0..w2, bytes r0 g0 b0 r1 ...) as
// three channel dwords, and back.
__device__ __forceinline__ void deinterleave4(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t &r, uint32_t &g,
                                              uint32_t &b) {
    r = (w0 & 0xffu) | ((w0 >> 24) & 0xffu) << 8 | ((w1 >> 16) & 0xffu) << 16 | ((w2 >> 8) & 0xffu) << 24;
    g = ((w0 >> 8) & 0xffu) | (w1 & 0xffu) << 8 | ((w1 >> 24) & 0xffu) << 16 | ((w2 >> 16) & 0xffu) << 24;
    b = ((w0 >> 16) & 0xffu) | ((w1 >> 8) & 0xffu) << 8 | (w2 & 0xffu) << 16 | ((w2 >> 24) & 0xffu) << 24;
}
__device__ __forceinline__ void interleave4(uint32_t r, uint32_t g, uint32_t b, uint32_t &w0, uint32_t &w1,
                                            uint32_t &w2) {
    w0 = (r & 0xffu) | (g & 0xffu) << 8 | (b & 0xffu) << 16 | ((r >> 8) & 0xffu) << 24;
    w1 = ((g >> 8) & 0xffu) | ((b >> 8) & 0xffu) << 8 | ((r >> 16) & 0xffu) << 16 | ((g >> 16) & 0xffu) << 24;
    w2 = ((b >> 16) & 0xffu) | ((r >> 24) & 0xffu) << 8 | ((g >> 24) & 0xffu) << 16 | ((b >> 24) & 0xffu) << 24;
}

__device__ __forceinline__ uint64_t cov_term(long long pix, uint32_t shape) {
    uint64_t z = ((static_cast<uint64_t>(pix) << 8) | shape) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ int32_t floor_div4(int32_t v) { return v >= 0 ? v / 4 : -((-v + 3) / 4); }

// Output rows a path draws (fill/mod.rs:176-203): documented mode draws device
// rows y_min/4..y_max/4; the literal loop draws y = aa_y/4 for aa_y == 4y+3 in
// [y_min, y_max].  Both are limited to 0 <= y < H.
__device__ __forceinline__ int2 drawn_rows(const PathHdr &h, int mode, int32_t height) {
    int32_t lo, hi;
    if (mode == RR_MODE_LITERAL) {
        lo = floor_div4(h.y_min);  // ceil((y_min - 3) / 4)
        hi = floor_div4(h.y_max - 3);
    } else {
        lo = h.y_min / 4;
        hi = h.y_max / 4;
    }
    return make_int2(max(lo, 0), min(hi, height - 1));
}


// ---------------------------------------------------------------------------
// kernel 2: scanline fill + composite of one band of 16 output rows
// ---------------------------------------------------------------------------
constexpr int kBandRows = 16;  // output rows per wave
constexpr int kFeatures = 12;  // dispatcher features per path
// Dispatcher MLP (12 -> 64 -> 64 -> 5, ReLU), parameters as one float array:
// feature mean and scale, W1 [64][12], b1, W2 [64][64], b2, W3 [5][64], b3.
constexpr int kHidden = 64, kClasses = 5;
constexpr int kMlpMu = 0, kMlpSd = kMlpMu + kFeatures, kMlpW1 = kMlpSd + kFeatures;
constexpr int kMlpB1 = kMlpW1 + kHidden * kFeatures, kMlpW2 = kMlpB1 + kHidden;
constexpr int kMlpB2 = kMlpW2 + kHidden * kHidden, kMlpW3 = kMlpB2 + kHidden;
constexpr int kMlpB3 = kMlpW3 + kClasses * kHidden, kMlpParams = kMlpB3 + kClasses;
// Tunable shape of a band fill (the autotuner picks among the instantiated
// variants): K = LDS entry slots per AA row (rows with more state-changing
// entries go to the wave's global slot), CH = pixels per composite sweep step
// (the LDS coverage bitmap holds 64 AA rows x 4*CH bits, padded by one dword
// per row against bank conflicts).
constexpr int kSlotK = 512;    // global overflow entry slots per AA row, per wave

template <int K, int CH>
struct BandLds {
    static constexpr int kWords = CH / 8;  // bitmap dwords per AA row
    static constexpr int kPitch = kWords + 1;
    int4 recA[1][64], recB[1][64];  // the chunk's EdgeRecs as their two 16-byte halves
    unsigned long long smask[64];  // per lane row: chunk edges whose row range starts or ends there
    int32_t ex0[K * 64];
    int32_t ex1[K * 64];
    int32_t ec[K * 64];
    uint32_t bm[64 * kPitch];
    uint8_t gamma[32];
};

struct Entry {
    int32_t x0, x1, c;
};

// Intersection of edge record q with AA row r (fill_intersections,
// scanner/mod.rs:528-585, with the per-row 16.16 stepping in closed form:
// the saturating adds of a constant step are absorbing, so step k equals the
// clamped 64-bit sum).
__device__ __forceinline__ Entry row_entry(const EdgeRec &q, int32_t r) {
    Entry en;
    const int32_t cnt = (q.meta & 4) ? 1 : -1;
    if (q.meta & EDGE_FAST) {
        // b + k*c stays inside int32 over the record's rows: 32-bit stepping,
        // clamped to [d, e] (d <= e; one v_med3_i32)
        const int32_t k = r - q.r0;
        const int32_t t1 = q.b + k * q.c;
        const int32_t px1 = min(max(t1, q.d), q.e) >> 16;
        const int32_t px0 = k == 0 ? q.a : min(max(t1 - q.c, q.d), q.e) >> 16;
        en.x0 = min(px0, px1);
        en.x1 = max(px0, px1);
        en.c = ((q.meta & 8) && k == 0) ? 0 : cnt;
        return en;
    }
    const int kind = q.meta & 3;
    if (kind == KIND_HORIZ) {
        en.x0 = q.a;
        en.x1 = q.b;
        en.c = 0;
        return en;
    }
    en.c = ((q.meta & 8) && r == q.r0) ? 0 : cnt;
    if (kind == KIND_VERT) {
        en.x0 = en.x1 = q.a;
        return en;
    }
    const long long k = static_cast<long long>(r) - q.r0;
    const long long t1 = static_cast<long long>(q.b) + k * q.c;
    int32_t v1 = sat64_to_i32(t1);
    v1 = v1 < q.d ? q.d : (v1 > q.e ? q.e : v1);
    const int32_t px1 = v1 >> 16;
    int32_t px0;
    if (k == 0) {
        px0 = q.a;
    } else {
        int32_t v0 = sat64_to_i32(t1 - q.c);
        v0 = v0 < q.d ? q.d : (v0 > q.e ? q.e : v0);
        px0 = v0 >> 16;
    }
    en.x0 = min(px0, px1);
    en.x1 = max(px0, px1);
    return en;
}

// Lane range [la, lb] of the band whose AA rows lie in [a, b].  Lane l holds
// AA row 4*Y0 + l (documented) or 16*(Y0 + l/4) + 12 + l%4 (literal).
__device__ __forceinline__ void lane_range(int32_t a, int32_t b, bool literal, int32_t Y0, int32_t band_lo,
                                           int32_t band_hi, int *la, int *lb) {
    if (!literal) {
        *la = max(a, band_lo) - band_lo;
        *lb = min(b, band_hi) - band_lo;
        return;
    }
    const int32_t base = 16 * Y0 + 12;
    int32_t rel = a - base;
    int lo_l;
    if (rel <= 0) {
        lo_l = 0;
    } else {
        const int32_t q = rel >> 4, m = rel & 15;
        lo_l = m <= 3 ? 4 * q + m : 4 * (q + 1);
    }
    rel = b - base;
    int hi_l;
    if (rel < 0) {
        hi_l = -1;
    } else {
        const int32_t q = rel >> 4, m = rel & 15;
        hi_l = m <= 3 ? 4 * q + m : 4 * q + 3;
    }
    *la = min(lo_l, 64);
    *lb = min(hi_l, 63);
}

// Walks the path's edges for this wave's band and builds each lane's list of
// retired, state-changing entries in X0/X1/C ([slot * 64 + lane]).  Chunks of
// 64 edges whose row summary misses the band are skipped with one ballot.
// Inside a chunk the edge->row relation is transposed: every edge marks the
// first and last lane it covers, a 64-bit prefix OR across the wave gives each
// lane the set of chunk edges crossing its row, and each lane folds only
// those, in edge (path) order.  Returns the number of entries this lane
// produced (may exceed cap: then only the first cap are stored and the caller
// reruns with more room).  sm0 is this lane's summary of chunk
// (e0 >> 6) + lane (chunk_summary_first), loaded by the caller ahead of time;
// X0/X1/C are LDS or global pointers (the two are separate instantiations).
__device__ __forceinline__ int2 chunk_summary_first(const FillArgs &A, int64_t e0, int64_t e1) {
    int2 sm = make_int2(INT32_MAX, INT32_MIN);
    const int lane = threadIdx.x & 63;
    if (e1 > e0 && (e0 >> 6) + lane <= ((e1 - 1) >> 6)) sm = A.csum[(e0 >> 6) + lane];
    return sm;
}

// Barrier of a band's wave: the whole (one-wave) workgroup, or with WS only
// this wave (a wave of a larger workgroup working alone: its LDS accesses
// execute in order, so a compiler fence suffices).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WS>
__device__ __forceinline__ void band_sync() {
    if constexpr (WS) {
        wave_sync();
    } else {
        __syncthreads();
    }
}

template <bool WS = false, typename Lds, typename Ptr>
__device__ __forceinline__ int build_entries(const FillArgs &A, Lds &L, int64_t e0, int64_t e1,
                                             int32_t band_lo, int32_t band_hi, bool literal, int32_t Y0, bool valid,
                                             int32_t r, int32_t mrule, Ptr X0, Ptr X1, Ptr C, int cap, int2 sm0) {
    const int lane = threadIdx.x & 63;
    int n = 0;
    // the row's open entry; the initial one (count 0, empty range) merges
    // into nothing it could change and retires as a no-op
    int32_t lx0 = INT32_MAX, lx1 = INT32_MIN, lc = 0;
    auto retire = [&](int32_t x0, int32_t x1, int32_t c) {
        if ((c & mrule) == 0) return;  // no effect on the walk
        if (n < cap) {
            X0[n * 64 + lane] = x0;
            X1[n * 64 + lane] = x1;
            C[n * 64 + lane] = c;
        }
        ++n;
    };
    // The chunk in LDS slot s (lane j's record in recA[s][j] / recB[s][j]):
    // fold the edges crossing this lane's row, in edge order.  r0/r1: this
    // lane's own record's rows.
    auto process = [&](int s, int32_t r0, int32_t r1) {
        int la = 64, lb = -1;
        if (r0 <= r1) lane_range(r0, r1, literal, Y0, band_lo, band_hi, &la, &lb);
        L.smask[lane] = 0ull;
        band_sync<WS>();
        // edge j toggles its bit on at lane la and off at lane lb + 1: the
        // prefix XOR leaves it set exactly on lanes la..lb.  Edges that start
        // at lane 0 (most: they enter the band from above) are a ballot
        // instead of a pile of atomics on one word.
        const unsigned long long from0 = __ballot(la == 0 && lb >= 0);
        if (la <= lb) {
            if (la > 0) atomicXor(&L.smask[la], 1ull << lane);
            if (lb < 63) atomicXor(&L.smask[lb + 1], 1ull << lane);
        }
        band_sync<WS>();
        unsigned long long m = wave_incl_xor64(L.smask[lane]) ^ from0;
        if (!valid) m = 0ull;
        while (m) {
            const int j = __ffsll(m) - 1;
            m &= m - 1;
            const int4 ga = L.recA[s][j], gb = L.recB[s][j];
            const EdgeRec g{ga.x, ga.y, ga.z, ga.w, gb.x, gb.y, gb.z, gb.w};
            const Entry en = row_entry(g, r);
            // write_intersect's test (saturating +1 on both sides)
            const bool touch = __builtin_elementwise_add_sat(lx1, 1) >= en.x0 &&
                               lx0 <= __builtin_elementwise_add_sat(en.x1, 1);
            if (!touch) retire(lx0, lx1, lc);
            lc = touch ? static_cast<int32_t>(static_cast<uint32_t>(lc) + static_cast<uint32_t>(en.c)) : en.c;
            lx0 = touch ? min(lx0, en.x0) : en.x0;
            lx1 = touch ? max(lx1, en.x1) : en.x1;
        }
        band_sync<WS>();
    };
    if (e1 <= e0) return 0;
    const int64_t c_first = e0 >> 6, c_last = (e1 - 1) >> 6;
    for (int64_t cb = c_first; cb <= c_last; cb += 64) {
        bool hit = false;
        if (cb + lane <= c_last) {
            const int2 sm = cb == c_first ? sm0 : A.csum[cb + lane];
            hit = sm.x <= band_hi && sm.y >= band_lo;
        }
        uint64_t cmask = __ballot(hit);
        // records of the overlapping chunks are loaded two chunks ahead
        auto load_rec = [&](int jc) {
            const int64_t i = ((cb + jc) << 6) + lane;
            EdgeRec q{1, 0, 0, 0, 0, 0, 0, 0};
            if (i >= e0 && i < e1) q = A.recs[i];
            return q;
        };
        EdgeRec qn{1, 0, 0, 0, 0, 0, 0, 0}, qnn{1, 0, 0, 0, 0, 0, 0, 0};
        if (cmask) qn = load_rec(__ffsll(static_cast<unsigned long long>(cmask)) - 1);
        {
            const uint64_t rest = cmask & (cmask - 1);
            if (rest) qnn = load_rec(__ffsll(static_cast<unsigned long long>(rest)) - 1);
        }
        while (cmask) {
            cmask &= cmask - 1;
            const EdgeRec q = qn;
            qn = qnn;
            const uint64_t after = cmask & (cmask - 1);
            if (after) qnn = load_rec(__ffsll(static_cast<unsigned long long>(after)) - 1);
            L.recA[0][lane] = make_int4(q.r0, q.r1, q.meta, q.a);
            L.recB[0][lane] = make_int4(q.b, q.c, q.d, q.e);
            process(0, q.r0, q.r1);
        }
    }
    retire(lx0, lx1, lc);
    return n;
}

// Stable insertion sort by x0, then the winding walk of render_aa_line.  The
// spans of the row are OR-ed into the AaBuf there; here they are merged into
// disjoint sorted intervals [s, e) clamped to [0, aa_width), written back in
// place over X0 (s) / X1 (e).  Returns the interval count; *unbalanced is set
// when the winding never returned to zero (the reference would loop forever
// there; the span runs to aa_width and the row stops).
template <typename Ptr>
__device__ __forceinline__ int walk_sorted(int n, Ptr X0, Ptr X1, Ptr C, int32_t mrule, int32_t aa_width,
                                           bool *unbalanced);

template <typename Ptr>
__device__ __forceinline__ int sort_and_walk(int n, Ptr X0, Ptr X1, Ptr C, int32_t mrule, int32_t aa_width,
                                             bool *unbalanced) {
    const int lane = threadIdx.x & 63;
    for (int a = 1; a < n; ++a) {
        const int32_t vx0 = X0[a * 64 + lane], vx1 = X1[a * 64 + lane], vc = C[a * 64 + lane];
        int b = a;
        while (b > 0) {
            const int32_t px = X0[(b - 1) * 64 + lane];
            if (px <= vx0) break;
            X0[b * 64 + lane] = px;
            X1[b * 64 + lane] = X1[(b - 1) * 64 + lane];
            C[b * 64 + lane] = C[(b - 1) * 64 + lane];
            --b;
        }
        if (b != a) {
            X0[b * 64 + lane] = vx0;
            X1[b * 64 + lane] = vx1;
            C[b * 64 + lane] = vc;
        }
    }
    return walk_sorted(n, X0, X1, C, mrule, aa_width, unbalanced);
}

// The winding walk of render_aa_line over entries already sorted by x0 (lane
// (threadIdx.x & 63)'s list, [slot * 64 + lane]): disjoint sorted intervals
// written back in place over X0 (s) / X1 (e); returns their count.
template <typename Ptr>
__device__ __forceinline__ int walk_sorted(int n, Ptr X0, Ptr X1, Ptr C, int32_t mrule, int32_t aa_width,
                                           bool *unbalanced) {
    const int lane = threadIdx.x & 63;
    int i = 0, nout = 0;
    int32_t cs = 0, ce = -1;
    int32_t count = 0;
    while (i < n) {
        // outside: every stored entry changes the state, so it opens a span
        const int32_t xx = X0[i * 64 + lane];
        count = static_cast<int32_t>(static_cast<uint32_t>(count) + static_cast<uint32_t>(C[i * 64 + lane]));
        ++i;
        int32_t end = aa_width;
        bool closed = false;
        while (i < n) {
            count = static_cast<int32_t>(static_cast<uint32_t>(count) + static_cast<uint32_t>(C[i * 64 + lane]));
            const int32_t x1 = X1[i * 64 + lane];
            ++i;
            if ((count & mrule) == 0) {
                end = x1 == INT32_MAX ? INT32_MAX : x1 + 1;
                closed = true;
                break;
            }
        }
        const int32_t bx0 = xx > 0 ? xx : 0;
        int32_t bx1 = end < aa_width ? end : aa_width;
        if (bx1 < 0) bx1 = 0;
        if (bx0 < bx1) {
            if (ce >= 0 && bx0 <= ce) {
                ce = max(ce, bx1);
            } else {
                if (ce >= 0) {
                    X0[nout * 64 + lane] = cs;
                    X1[nout * 64 + lane] = ce;
                    ++nout;
                }
                cs = bx0;
                ce = bx1;
            }
        }
        if (!closed) {
            *unbalanced = true;
            break;
        }
    }
    if (ce >= 0) {
        X0[nout * 64 + lane] = cs;
        X1[nout * 64 + lane] = ce;
        ++nout;
    }
    return nout;
}


}  // namespace rrdev
