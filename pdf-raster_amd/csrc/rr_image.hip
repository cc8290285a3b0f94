// rr_image.hip -- gfx950 kernels for image placement + soft mask + Normal /
// Multiply compositing through a grey CLUT, and the reference's elementwise
// GPU ops (ICC CLUT / matrix, RGBA composite, soft mask), bit-exact to the
// reference CPU twins.
//
// Image placement follows the reference's general affine path of blit_image
// (crates/pdf_interp/src/renderer/page/mod.rs:1607-1676 bbox + inverse CTM,
// :1899-1974 per-pixel sampling): every canvas pixel of the clamped bounding
// box maps back through the inverse CTM to a clamped (u, v), hence to an image
// pixel and its soft-mask byte; mask 0 skips, 255 copies, anything else is
// blend_u8 (:2146-2149).  The colour goes to grey through a u16 grid (3-D for
// RGB, 4-D for CMYK) with integer simplex interpolation, so the kernel and the
// CPU oracle agree bit for bit.
//
// Work decomposition: one launch per image (images of a page are ordered) of
// persistent 1024-thread blocks.  Each block first copies the grid (or, when it
// is larger than the block's LDS, its first nodes) into LDS, then walks
// 256 px x 16-row canvas tiles: a wave per 32 px x 8-row block, one thread per
// 4 consecutive canvas pixels so every canvas access is one aligned dword
// read-modify-write.  Grid taps read LDS, or global memory past the LDS copy.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "rr_hip.h"
#include "rr_internal.h"

namespace {

struct ImgArgs {
    const uint8_t *pixels;
    const uint8_t *smask;
    uint8_t *canvas;
    const uint16_t *grid;
    uint32_t g;
    int32_t channels, blend, img_w;
    int32_t stride, bx0, bx1, by0, by1;
    uint32_t lds_nodes;      // grid nodes [0, lds_nodes) are read from the LDS copy
    uint32_t last_node;      // g^D - 1
    int32_t tiles_x, n_tiles;  // 256 px x 16-row canvas tiles over the bbox
    double a, b, c, d, e, f, inv_det, page_h, iw, ih;
    int64_t n_img_px;          // width * height of the image
    const uint16_t *grey;      // grey | alpha << 8 per image pixel, when converted ahead
    uint16_t *grey_out;        // where k_image_grey writes them
    // axis-aligned path (|b| w < 0.5 and |c| h < 0.5): image x of canvas column
    // bx0 + k is (ix_fp0 + k * ix_step) >> 32 in Q32, wrapping, clamped
    int32_t axis;
    int64_t ix_fp0, ix_step;
};

__device__ __forceinline__ uint32_t div255(uint32_t x) {
    const uint32_t s = (x + (x >> 8) + 0x80u) >> 8;
    return s < 255u ? s : 255u;
}

// 24-bit multiplies (full-rate VALU).  Operands must be below 2^24; the
// compiler otherwise picks the 32-bit multiply for these products.
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// Integer simplex interpolation of a grey u16 grid: for channel bytes c,
// t = c*(g-1) splits into a cell index and a fraction in 1/255 units; the axes,
// taken in decreasing fraction order (equal fractions keep channel order), step
// from the cell's base node to its far corner.  Returns the D + 1 nodes and
// their weights (summing to 255); the grey value is
// (sum of weight * node + 32767) / 65535.
//
// The order comes from a sorting network over keys fraction * 4 + (3 - axis):
// distinct keys whose descending order is exactly decreasing fraction with
// ties in channel order; each key carries its axis' stride along.  An axis at
// the grid's last node (byte 255) has fraction 0, so it sorts behind every
// axis with a nonzero fraction and every vertex that steps along it has
// weight 0: such a vertex's node is clamped into the grid (last_node) instead
// of holding the step back, which leaves the grey value unchanged.
template <int D>
__device__ __forceinline__ void clut_cell(const uint32_t c[D], uint32_t g, uint32_t last_node, uint32_t idx[D + 1],
                                          uint32_t w[D + 1]) {
    uint32_t key[D], st[D], i0[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        // t < 2^16 and g <= 256, so 24-bit multiplies are exact, and
        // (t * 32897) >> 23 == t / 255 for every t < 2^16
        const uint32_t t = mul24(c[a], g - 1u);
        i0[a] = mul24(t, 32897u) >> 23;
        const uint32_t fr = t - mul24(i0[a], 255u);
        key[a] = fr << 2 | static_cast<uint32_t>(3 - a);
    }
    st[0] = g * g;  // channel 0 outermost of the first three axes; K outermost of all
    st[1] = g;
    st[2] = 1u;
    if constexpr (D == 4) st[3] = g * g * g;
    // base node by Horner's rule over (K,) channel 0, 1, 2: every partial
    // product stays below 2^32 with operands below 2^24
    uint32_t base = D == 4 ? mad24(i0[D - 1], g, i0[0]) : i0[0];
    base = mad24(base, g, i0[1]);
    base = mad24(base, g, i0[2]);
    auto cs = [&](int x, int y) {
        const bool sw = key[y] > key[x];
        const uint32_t hi = max(key[x], key[y]), lo = min(key[x], key[y]);
        const uint32_t sx = sw ? st[y] : st[x], sy = sw ? st[x] : st[y];
        key[x] = hi;
        key[y] = lo;
        st[x] = sx;
        st[y] = sy;
    };
    if constexpr (D == 4) {
        cs(0, 1);
        cs(2, 3);
        cs(0, 2);
        cs(1, 3);
        cs(1, 2);
    } else {
        cs(0, 1);
        cs(1, 2);
        cs(0, 1);
    }
    idx[0] = base;
    w[0] = 255u - (key[0] >> 2);
    uint32_t cur = base;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        cur += st[k];
        idx[k + 1] = min(cur, last_node);
        w[k + 1] = (key[k] >> 2) - (k + 1 < D ? key[k + 1] >> 2 : 0u);
    }
}

constexpr int kImgThreads = 1024;  // 16 waves, each on a 32 px x 8-row block
constexpr int kImgTileW = 256;     // tile: 8 x 2 wave blocks
constexpr int kImgTileH = 16;

// Colour bytes and soft-mask byte of image pixel idx.
template <int D>
__device__ __forceinline__ void load_pixel(const ImgArgs &A, int64_t idx, uint32_t c[D], uint32_t &alpha) {
    alpha = A.smask ? A.smask[idx] : 255u;
    if constexpr (D == 4) {
        const uint32_t q = *reinterpret_cast<const uint32_t *>(A.pixels + idx * 4);
#pragma unroll
        for (int a = 0; a < 4; ++a) c[a] = (q >> (8 * a)) & 0xffu;
    } else {
        // the 3 bytes as one aligned 16-bit load and one byte load, both
        // inside the pixel
        const uintptr_t p = reinterpret_cast<uintptr_t>(A.pixels) + static_cast<uintptr_t>(idx) * 3;
        const bool odd = p & 1u;
        const uint32_t h = *reinterpret_cast<const uint16_t *>(odd ? p + 1 : p);
        const uint32_t b = *reinterpret_cast<const uint8_t *>(odd ? p : p + 2);
        const uint32_t q = odd ? (b | h << 8) : (h | b << 16);
#pragma unroll
        for (int a = 0; a < 3; ++a) c[a] = (q >> (8 * a)) & 0xffu;
    }
}

// Grey values of 4 colours through the grid (the simplex taps, then the
// rounded division), taps from the LDS copy when all of a lane's lie in it.
template <int D>
__device__ __forceinline__ void grey4(const ImgArgs &A, const uint16_t *lgrid, const uint32_t c[4][D], uint32_t s[4]) {
    uint32_t tap[4][D + 1], wt[4][D + 1], gi[4][D + 1];
    uint32_t gmax = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        clut_cell<D>(c[k], A.g, A.last_node, gi[k], wt[k]);
#pragma unroll
        for (int v = 0; v <= D; ++v) gmax = max(gmax, gi[k][v]);
    }
    // a lane whose taps all lie in the LDS copy reads them there; any other
    // lane reads all of its taps from the grid in global memory (one branch,
    // so each side's loads are in flight together)
    if (gmax < A.lds_nodes) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int v = 0; v <= D; ++v) tap[k][v] = lgrid[gi[k][v]];
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int v = 0; v <= D; ++v) tap[k][v] = A.grid[gi[k][v]];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t acc = 0;
#pragma unroll
        for (int v = 0; v <= D; ++v) acc = mad24(wt[k][v], tap[k][v], acc);
        // x / 65535 == (x + 1 + (x >> 16)) >> 16 for x < 2^24
        const uint32_t x = acc + 32767u;
        s[k] = (x + 1u + (x >> 16)) >> 16;
    }
}

// One thread per aligned canvas dword (4 pixels).  The work runs in phases
// over the 4 pixels, so that each phase's loads are in flight together:
// sample positions, then the sampled image pixels' colour and soft-mask bytes
// (or their grey + alpha pairs when the image was converted ahead, GREY),
// then the simplex grid taps, then the blends (blend_multiply, blend_u8).
template <int D, bool GREY>
__device__ __forceinline__ void place_dword(const ImgArgs &A, const uint16_t *lgrid, int32_t x0, int32_t dy) {
    const double dy_pdf = A.page_h - static_cast<double>(dy);
    const double dy_rel = dy_pdf - A.f;
    int64_t idx[4];
    bool in[4];
    if (A.axis) {
        // page/mod.rs:1811-1829: v and iy per row in f64; the Q32 column
        // stepping of :1819-1833 in closed form (wrapping, as the reference's
        // wrapping_add per column)
        double v = dy_rel / A.d;
        v = fmin(fmax(v, 0.0), 1.0);  // a NaN v ends as iy = ih - 1 (f64::min drops the NaN)
        const int64_t row = static_cast<int64_t>(static_cast<int32_t>(fmin((1.0 - v) * A.ih, A.ih - 1.0))) * A.img_w;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t dx = x0 + k;
            in[k] = dx >= A.bx0 && dx < A.bx1;
            const uint64_t fp = static_cast<uint64_t>(A.ix_fp0) +
                                static_cast<uint64_t>(static_cast<int64_t>(dx - A.bx0)) * static_cast<uint64_t>(A.ix_step);
            int64_t ix = static_cast<int64_t>(fp) >> 32;
            ix = ix < 0 ? 0 : (ix > A.img_w - 1 ? A.img_w - 1 : ix);
            idx[k] = in[k] ? row + ix : 0;
        }
    } else {
        const double u_row = (-A.c * dy_rel) * A.inv_det;
        const double v_row = (A.a * dy_rel) * A.inv_det;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t dx = x0 + k;
            in[k] = dx >= A.bx0 && dx < A.bx1;
            const double dx_rel = static_cast<double>(dx) - A.e;
            double u = fma(A.d * dx_rel, A.inv_det, u_row);
            double v = fma((-A.b) * dx_rel, A.inv_det, v_row);
            // clamps to [0, 1]; a NaN u or v (from overflowing products) ends as
            // ix = iw - 1, iy = ih - 1, as it does when the clamp keeps the NaN
            u = fmax(fmin(u, 1.0), 0.0);
            v = fmin(fmax(v, 0.0), 1.0);
            const int32_t ix = static_cast<int32_t>(fmin(u * A.iw, A.iw - 1.0));
            const int32_t iy = static_cast<int32_t>(fmin((1.0 - v) * A.ih, A.ih - 1.0));
            // a pixel outside the bbox reads image pixel 0 and is discarded
            idx[k] = in[k] ? static_cast<int64_t>(iy) * A.img_w + ix : 0;
        }
    }
    uint32_t s[4], alpha[4];
    if constexpr (GREY) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t q = A.grey[idx[k]];
            s[k] = q & 0xffu;
            alpha[k] = q >> 8;
        }
    } else {
        uint32_t c[4][D];
#pragma unroll
        for (int k = 0; k < 4; ++k) load_pixel<D>(A, idx[k], c[k], alpha[k]);
        grey4<D>(A, lgrid, c, s);
    }
    uint32_t *dst = reinterpret_cast<uint32_t *>(A.canvas + static_cast<int64_t>(dy) * A.stride + x0);
    const uint32_t old = *dst;
    uint32_t out = old;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (!in[k] || alpha[k] == 0) continue;
        any = true;
        uint32_t sk = s[k];
        const uint32_t dv = (old >> (8 * k)) & 0xffu;
        // blend_multiply.  The product goes through mul24: with both factors
        // known to be bytes, the compiler folds the plain div255(sk * dv)
        // into a v_dot4_u32_u8 sequence whose results differ from the
        // formula's (seen on enlarged images, where sk is a loaded byte).
        if (A.blend == 1) sk = div255(mul24(sk, dv));
        const uint32_t al = alpha[k];
        const uint32_t r = al == 255u ? sk : mul24(sk * al + dv * (255u - al), 32897u) >> 23;  // blend_u8, / 255
        out = (out & ~(0xffu << (8 * k))) | (r << (8 * k));
    }
    if (any) *dst = out;
}

// Copy grid nodes [0, lds_nodes) into LDS (16-byte loads when the grid is
// 16-byte aligned).
__device__ __forceinline__ void grid_to_lds(const ImgArgs &A, uint4 *lds_raw) {
    uint16_t *lgrid = reinterpret_cast<uint16_t *>(lds_raw);
    const int32_t tid = static_cast<int32_t>(threadIdx.x);
    int32_t done = 0;
    if ((reinterpret_cast<uintptr_t>(A.grid) & 15u) == 0) {
        const int32_t n16 = A.lds_nodes >> 3;
        const uint4 *src = reinterpret_cast<const uint4 *>(A.grid);
#pragma unroll 4
        for (int32_t i = tid; i < n16; i += kImgThreads) lds_raw[i] = src[i];
        done = n16 << 3;
    }
    for (int32_t i = done + tid; i < static_cast<int32_t>(A.lds_nodes); i += kImgThreads) lgrid[i] = A.grid[i];
    __syncthreads();
}

// Persistent blocks: the grid into LDS (unless the image was converted ahead),
// then tiles blockIdx.x, blockIdx.x + gridDim.x, ...
template <int D, bool GREY>
__global__ void __launch_bounds__(kImgThreads) k_place_image(ImgArgs A) {
    extern __shared__ uint4 lds_raw[];
    if constexpr (!GREY) grid_to_lds(A, lds_raw);
    const uint16_t *lgrid = reinterpret_cast<const uint16_t *>(lds_raw);
    const int32_t tid = static_cast<int32_t>(threadIdx.x);
    // a wave covers a 32 px x 8-row block (8 x 8 lanes), so that under rotation
    // its image samples fall in a compact patch of few cache lines
    const int32_t xs = A.bx0 & ~3;
    const int32_t wave = tid >> 6, lane = tid & 63;
    const int32_t ox = (wave & 7) * 32 + (lane & 7) * 4, oy = (wave >> 3) * 8 + (lane >> 3);
    for (int32_t t = static_cast<int32_t>(blockIdx.x); t < A.n_tiles; t += static_cast<int32_t>(gridDim.x)) {
        const int32_t ty = t / A.tiles_x, tx = t - ty * A.tiles_x;
        const int32_t x0 = xs + tx * kImgTileW + ox;
        const int32_t dy = A.by0 + ty * kImgTileH + oy;
        if (dy < A.by1 && x0 < A.bx1) place_dword<D, GREY>(A, lgrid, x0, dy);
    }
}

// Grey + alpha of every image pixel ahead of an enlarging placement, 4 pixels
// a thread: grey[i] = grey value | soft-mask byte << 8 (the values the
// placement would compute for each canvas pixel sampling pixel i).
template <int D>
__global__ void __launch_bounds__(kImgThreads) k_image_grey(ImgArgs A) {
    extern __shared__ uint4 lds_raw[];
    grid_to_lds(A, lds_raw);
    const uint16_t *lgrid = reinterpret_cast<const uint16_t *>(lds_raw);
    const int64_t n = A.n_img_px;
    for (int64_t g = static_cast<int64_t>(blockIdx.x) * kImgThreads + threadIdx.x; 4 * g < n;
         g += static_cast<int64_t>(gridDim.x) * kImgThreads) {
        uint32_t c[4][D], alpha[4], s[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) load_pixel<D>(A, min(4 * g + k, n - 1), c[k], alpha[k]);
        grey4<D>(A, lgrid, c, s);
        uint16_t *out = A.grey_out + 4 * g;
        if (4 * g + 4 <= n) {
            *reinterpret_cast<uint2 *>(out) = make_uint2(s[0] | alpha[0] << 8 | (s[1] | alpha[1] << 8) << 16,
                                                         s[2] | alpha[2] << 8 | (s[3] | alpha[3] << 8) << 16);
        } else {
            for (int k = 0; 4 * g + k < n; ++k) out[k] = static_cast<uint16_t>(s[k] | alpha[k] << 8);
        }
    }
}

// icc_cmyk_to_rgb_cpu, CLUT arm (crates/gpu/src/cmyk.rs:531-630): f32
// quadrilinear with fma lerps in the reference's order; round half away.
__global__ void __launch_bounds__(256) k_icc_clut(const uint8_t *__restrict__ cmyk, uint8_t *__restrict__ rgb,
                                                  const uint8_t *__restrict__ table, int32_t g, float scale,
                                                  int64_t n) {
    const int64_t g2 = static_cast<int64_t>(g) * g, g3 = g2 * g;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uchar4 s = reinterpret_cast<const uchar4 *>(cmyk)[i];
        const float fc = static_cast<float>(s.x) * scale, fm = static_cast<float>(s.y) * scale;
        const float fy = static_cast<float>(s.z) * scale, fk = static_cast<float>(s.w) * scale;
        const int ic0 = min(static_cast<int>(fc), g - 1), im0 = min(static_cast<int>(fm), g - 1);
        const int iy0 = min(static_cast<int>(fy), g - 1), ik0 = min(static_cast<int>(fk), g - 1);
        const int ic1 = min(ic0 + 1, g - 1), im1 = min(im0 + 1, g - 1), iy1 = min(iy0 + 1, g - 1), ik1 = min(ik0 + 1, g - 1);
        const float wc = fc - static_cast<float>(ic0), wm = fm - static_cast<float>(im0);
        const float wy = fy - static_cast<float>(iy0), wk = fk - static_cast<float>(ik0);
        auto lerp = [](float a, float b, float t) { return fmaf(t, b - a, a); };
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            auto node = [&](int ci, int mi, int yi, int ki) {
                return static_cast<float>(table[(ki * g3 + ci * g2 + static_cast<int64_t>(mi) * g + yi) * 3 + ch]);
            };
            const float c0m0k0 = lerp(node(ic0, im0, iy0, ik0), node(ic0, im0, iy1, ik0), wy);
            const float c0m1k0 = lerp(node(ic0, im1, iy0, ik0), node(ic0, im1, iy1, ik0), wy);
            const float c1m0k0 = lerp(node(ic1, im0, iy0, ik0), node(ic1, im0, iy1, ik0), wy);
            const float c1m1k0 = lerp(node(ic1, im1, iy0, ik0), node(ic1, im1, iy1, ik0), wy);
            const float rk0 = lerp(lerp(c0m0k0, c0m1k0, wm), lerp(c1m0k0, c1m1k0, wm), wc);
            const float c0m0k1 = lerp(node(ic0, im0, iy0, ik1), node(ic0, im0, iy1, ik1), wy);
            const float c0m1k1 = lerp(node(ic0, im1, iy0, ik1), node(ic0, im1, iy1, ik1), wy);
            const float c1m0k1 = lerp(node(ic1, im0, iy0, ik1), node(ic1, im0, iy1, ik1), wy);
            const float c1m1k1 = lerp(node(ic1, im1, iy0, ik1), node(ic1, im1, iy1, ik1), wy);
            const float rk1 = lerp(lerp(c0m0k1, c0m1k1, wm), lerp(c1m0k1, c1m1k1, wm), wc);
            float out = lerp(rk0, rk1, wk);
            out = out < 0.0f ? 0.0f : (out > 255.0f ? 255.0f : out);
            rgb[3 * i + ch] = static_cast<uint8_t>(roundf(out));
        }
    }
}

// cmyk_to_rgb_reflectance (crates/color/src/convert.rs:131-161)
__global__ void __launch_bounds__(256) k_cmyk_matrix(const uint8_t *__restrict__ cmyk, uint8_t *__restrict__ rgb,
                                                     int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uchar4 s = reinterpret_cast<const uchar4 *>(cmyk)[i];
        const uint32_t inv_k = 255u - s.w;
        rgb[3 * i] = static_cast<uint8_t>(((255u - s.x) * inv_k + 127u) / 255u);
        rgb[3 * i + 1] = static_cast<uint8_t>(((255u - s.y) * inv_k + 127u) / 255u);
        rgb[3 * i + 2] = static_cast<uint8_t>(((255u - s.z) * inv_k + 127u) / 255u);
    }
}

// composite_rgba8_cpu (crates/gpu/src/composite.rs:4-37)
__global__ void __launch_bounds__(256) k_composite_rgba8(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst,
                                                         int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const uchar4 s = reinterpret_cast<const uchar4 *>(src)[i];
        const uint32_t a_src = s.w;
        if (a_src == 0) continue;
        uchar4 *dp = reinterpret_cast<uchar4 *>(dst) + i;
        if (a_src == 255) {
            *dp = s;
            continue;
        }
        const uchar4 d = *dp;
        const uint32_t a_dst = d.w, inv = 255u - a_src;
        const uint32_t a_out = a_src + (a_dst * inv + 127u) / 255u;
        if (a_out == 0) continue;
        auto ch = [&](uint32_t sc, uint32_t dc) {
            const uint32_t b = (sc * a_src + dc * a_dst * inv / 255u + a_out / 2u) / a_out;
            return static_cast<unsigned char>(b < 255u ? b : 255u);
        };
        *dp = make_uchar4(ch(s.x, d.x), ch(s.y, d.y), ch(s.z, d.z), static_cast<unsigned char>(a_out));
    }
}

// apply_soft_mask_cpu (crates/gpu/src/composite.rs:39-52)
__global__ void __launch_bounds__(256) k_apply_soft_mask(uint8_t *__restrict__ px, const uint8_t *__restrict__ mask,
                                                         int64_t n) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        px[4 * i + 3] = static_cast<uint8_t>((static_cast<uint32_t>(px[4 * i + 3]) * mask[i] + 127u) / 255u);
}

// rgb_to_gray (crates/pdf_raster/src/render.rs:1306-1324): BT.709 integer
// weights, (2126 r + 7152 g + 722 b + 5000) / 10000.  One thread per 4 pixels
// of a row: three dword loads when the row's 12-byte groups are 4-byte aligned
// (aligned == 1), byte loads otherwise; one dword store when the grey row is.
__global__ void __launch_bounds__(256) k_rgb_to_gray(const uint8_t *__restrict__ rgb, int64_t rgb_stride,
                                                     uint8_t *__restrict__ gray, int64_t gray_stride, int32_t width,
                                                     int32_t height, int32_t aligned) {
    const int32_t x0 = 4 * (static_cast<int32_t>(blockIdx.x) * 256 + static_cast<int32_t>(threadIdx.x));
    const int32_t y = static_cast<int32_t>(blockIdx.y);
    if (x0 >= width || y >= height) return;
    const uint8_t *src = rgb + y * rgb_stride + 3 * static_cast<int64_t>(x0);
    uint8_t *dst = gray + y * gray_stride + x0;
    const int32_t n = min(4, width - x0);
    uint32_t c[12];
    if (aligned && n == 4) {
        const uint32_t *s4 = reinterpret_cast<const uint32_t *>(src);
        const uint32_t w0 = s4[0], w1 = s4[1], w2 = s4[2];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            c[k] = (w0 >> (8 * k)) & 0xffu;
            c[4 + k] = (w1 >> (8 * k)) & 0xffu;
            c[8 + k] = (w2 >> (8 * k)) & 0xffu;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) c[k] = k < 3 * n ? src[k] : 0u;
    }
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t g = (2126u * c[3 * k] + 7152u * c[3 * k + 1] + 722u * c[3 * k + 2] + 5000u) / 10000u;
        out |= g << (8 * k);
    }
    if (aligned && n == 4) {
        *reinterpret_cast<uint32_t *>(dst) = out;
    } else {
        for (int k = 0; k < n; ++k) dst[k] = static_cast<uint8_t>(out >> (8 * k));
    }
}

#define RR_TRY(ctx, expr)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return rrint::fail(ctx, RR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// f64 -> i64 as Rust's `as i64` converts: truncation, saturating, NaN -> 0.
int64_t sat_i64(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return static_cast<int64_t>(x);
}

unsigned grid_for(int64_t n) {
    const int64_t blocks = (n + 255) / 256;
    return static_cast<unsigned>(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
}

template <typename F>
int32_t launch_profiled(rr_hip_ctx *c, F &&launch) {
    bool ok = true;
    hipEvent_t *ev = rrint::prof_pair(c, &ok);
    if (!ok) return rrint::fail(c, RR_E_DEVICE, "profiling event allocation failed");
    if (ev) RR_TRY(c, hipEventRecord(ev[0], rrint::stream(c)));
    launch();
    RR_TRY(c, hipGetLastError());
    if (ev) RR_TRY(c, hipEventRecord(ev[1], rrint::stream(c)));
    rrint::count_launch(c);
    return RR_OK;
}

}  // namespace

namespace rrint {
int32_t rgb_to_gray(rr_hip_ctx *c, const uint8_t *rgb, int64_t rgb_stride, uint8_t *gray, int64_t gray_stride,
                    int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return RR_OK;
    if (height > 65535) return fail(c, RR_E_INVARIANT, "rgb_to_gray: height > 65535");
    const int32_t aligned = ((reinterpret_cast<uintptr_t>(rgb) | reinterpret_cast<uintptr_t>(gray) |
                              static_cast<uintptr_t>(rgb_stride) | static_cast<uintptr_t>(gray_stride)) & 3u) == 0;
    const dim3 grid(static_cast<unsigned>((width + 1023) / 1024), static_cast<unsigned>(height));
    return launch_profiled(c, [&] {
        hipLaunchKernelGGL(k_rgb_to_gray, grid, dim3(256), 0, stream(c), rgb, rgb_stride, gray, gray_stride, width,
                           height, aligned);
    });
}
}  // namespace rrint

extern "C" {

int32_t rr_hip_record_place_images(rr_hip_ctx *c, const rr_image_batch *b) {
    if (!c || !b) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (b->n_images < 0 || b->n_pages < 0 || (b->n_images > 0 && (!b->images || !b->pages)))
        return rrint::fail(c, RR_E_INVARIANT, "image batch: bad image/page arrays");
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    int lds_max = 0;
    RR_TRY(c, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, rrint::device(c)));
    // Pass 1: every descriptor is checked and its placement set up before
    // anything is launched, so a rejected batch leaves every canvas as it was
    // (a caller falling back to the CPU then applies each image exactly once).
    std::vector<ImgArgs> plan;
    plan.reserve(static_cast<size_t>(b->n_images));
    for (int32_t i = 0; i < b->n_images; ++i) {
        const rr_image_desc &im = b->images[i];
        if (im.page < 0 || im.page >= b->n_pages) return rrint::fail(c, RR_E_INVARIANT, "image: bad page index");
        const rr_page_desc &pg = b->pages[im.page];
        if (!pg.canvas || (reinterpret_cast<uintptr_t>(pg.canvas) & 3u) || pg.width <= 0 || pg.height <= 0 ||
            pg.stride < ((pg.width + 3) & ~3) || (pg.stride & 3) || pg.height > 4 * 65535)
            return rrint::fail(c, RR_E_INVARIANT, "image: bad canvas geometry");
        if (im.width <= 0 || im.height <= 0 || !im.pixels || (im.channels != 3 && im.channels != 4) ||
            (im.blend != 0 && im.blend != 1))
            return rrint::fail(c, RR_E_INVARIANT, "image: bad descriptor");
        const uint16_t *grid = im.channels == 4 ? b->grid4 : b->grid3;
        const int32_t g = im.channels == 4 ? b->grid4_n : b->grid3_n;
        if (!grid || g < 2 || g > 256) return rrint::fail(c, RR_E_INVARIANT, "image: missing or bad colour grid");
        if (im.channels == 4 && (reinterpret_cast<uintptr_t>(im.pixels) & 3u))
            return rrint::fail(c, RR_E_INVARIANT, "image: CMYK pixels must be 4-byte aligned");
        // bbox and inverse CTM on the host, in the reference's f64 operations
        const double *m = im.ctm;
        const double page_h = static_cast<double>(pg.height);
        const double ux[4] = {0.0, 1.0, 0.0, 1.0}, uy[4] = {0.0, 0.0, 1.0, 1.0};
        double xs[4], ys[4];
        bool finite = true;
        for (int k = 0; k < 4; ++k) {
            xs[k] = std::fma(m[0], ux[k], m[2] * uy[k]) + m[4];
            ys[k] = std::fma(m[1], ux[k], m[3] * uy[k]) + m[5];
            finite = finite && std::isfinite(xs[k]) && std::isfinite(ys[k]);
        }
        if (!finite) continue;  // reference: warn and skip the image
        const double xmn = std::fmin(std::fmin(xs[0], xs[1]), std::fmin(xs[2], xs[3]));
        const double xmx = std::fmax(std::fmax(xs[0], xs[1]), std::fmax(xs[2], xs[3]));
        const double ymn = std::fmin(std::fmin(page_h - ys[0], page_h - ys[1]), std::fmin(page_h - ys[2], page_h - ys[3]));
        const double ymx = std::fmax(std::fmax(page_h - ys[0], page_h - ys[1]), std::fmax(page_h - ys[2], page_h - ys[3]));
        auto clampi = [](int64_t v, int64_t hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
        const int64_t bx0 = clampi(sat_i64(std::floor(xmn)), pg.width);
        const int64_t bx1 = clampi(sat_i64(std::ceil(xmx)), pg.width);
        const int64_t by0 = clampi(sat_i64(std::floor(ymn)), pg.height);
        const int64_t by1 = clampi(sat_i64(std::ceil(ymx)), pg.height);
        if (bx0 >= bx1 || by0 >= by1) continue;
        const double det = std::fma(m[0], m[3], -(m[1] * m[2]));
        if (std::fabs(det) < 1e-12) continue;
        const double iw = static_cast<double>(im.width), ih = static_cast<double>(im.height);
        ImgArgs A{};
        A.pixels = im.pixels;
        A.smask = im.smask;
        A.canvas = pg.canvas;
        A.grid = grid;
        A.g = static_cast<uint32_t>(g);
        A.channels = im.channels;
        A.blend = im.blend;
        A.img_w = im.width;
        A.stride = pg.stride;
        A.bx0 = static_cast<int32_t>(bx0);
        A.bx1 = static_cast<int32_t>(bx1);
        A.by0 = static_cast<int32_t>(by0);
        A.by1 = static_cast<int32_t>(by1);
        A.a = m[0];
        A.b = m[1];
        A.c = m[2];
        A.d = m[3];
        A.e = m[4];
        A.f = m[5];
        A.inv_det = 1.0 / det;
        A.page_h = page_h;
        A.iw = iw;
        A.ih = ih;
        // axis-aligned path (page/mod.rs:1767-1809): the Q32 column step
        // img_w / a and the start column's image x, as Rust's `as i64`
        A.axis = std::fabs(m[1]) * iw < 0.5 && std::fabs(m[2]) * ih < 0.5;
        if (A.axis) {
            const double fp_scale = 4294967296.0;
            A.ix_step = sat_i64((iw / m[0]) * fp_scale);
            double u0 = (static_cast<double>(bx0) - m[4]) / m[0];
            u0 = u0 < 0.0 ? 0.0 : (u0 > 1.0 ? 1.0 : u0);  // f64::clamp keeps a NaN
            A.ix_fp0 = sat_i64(u0 * iw * fp_scale);
        }
        const int32_t xs0 = A.bx0 & ~3;
        A.tiles_x = (A.bx1 - xs0 + kImgTileW - 1) / kImgTileW;
        A.n_tiles = A.tiles_x * ((A.by1 - A.by0 + kImgTileH - 1) / kImgTileH);
        // as much of the grid as one block's LDS holds; the rest stays in global memory
        int64_t nodes = g;
        for (int k = 1; k < im.channels; ++k) nodes *= g;
        const int64_t lds_nodes = std::min<int64_t>(nodes, (lds_max / 2) & ~int64_t{7});
        if (lds_nodes < 8) return rrint::fail(c, RR_E_DEVICE, "image: device has no LDS for the grid");
        A.lds_nodes = static_cast<uint32_t>(lds_nodes);
        A.last_node = static_cast<uint32_t>(nodes - 1);
        A.n_img_px = static_cast<int64_t>(im.width) * im.height;
        plan.push_back(A);
    }
    // Pass 2: the launches, in image order.
    for (ImgArgs &A : plan) {
        const size_t lds_bytes = static_cast<size_t>(A.lds_nodes) * 2;
        // an image smaller than 0.8 of its canvas footprint is converted to
        // grey + alpha once per image pixel, so that the placement samples
        // those instead of running the grid interpolation per canvas pixel
        A.grey = nullptr;
        A.grey_out = nullptr;
        const int64_t n_box = static_cast<int64_t>(A.bx1 - A.bx0) * (A.by1 - A.by0);
        const bool ahead = A.n_img_px * 5 < n_box * 4;
        auto blocks_of = [&](const void *kfn, size_t lds, int64_t work, int64_t *blocks) -> int32_t {
            RR_TRY(c, hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
            int per_cu = 0;
            RR_TRY(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kImgThreads, lds));
            *blocks = std::max<int64_t>(1, std::min<int64_t>(work, static_cast<int64_t>(c->n_cu) * std::max(1, per_cu)));
            return RR_OK;
        };
        int64_t blocks = 0;
        int32_t rc = RR_OK;
        if (ahead) {
            rc = rrint::ensure(c, &c->img_grey, &c->img_grey_cap, static_cast<size_t>(A.n_img_px));
            if (rc) return rc;
            A.grey_out = c->img_grey;
            const void *kg = A.channels == 4 ? reinterpret_cast<const void *>(&k_image_grey<4>)
                                             : reinterpret_cast<const void *>(&k_image_grey<3>);
            rc = blocks_of(kg, lds_bytes, (A.n_img_px + 4 * kImgThreads - 1) / (4 * kImgThreads), &blocks);
            if (rc) return rc;
            rc = launch_profiled(c, [&] {
                if (A.channels == 4)
                    hipLaunchKernelGGL(k_image_grey<4>, dim3(static_cast<unsigned>(blocks)), dim3(kImgThreads),
                                       lds_bytes, rrint::stream(c), A);
                else
                    hipLaunchKernelGGL(k_image_grey<3>, dim3(static_cast<unsigned>(blocks)), dim3(kImgThreads),
                                       lds_bytes, rrint::stream(c), A);
            });
            if (rc) return rc;
            A.grey = c->img_grey;
            rc = blocks_of(reinterpret_cast<const void *>(&k_place_image<4, true>), 0, A.n_tiles, &blocks);
            if (rc) return rc;
            rc = launch_profiled(c, [&] {
                hipLaunchKernelGGL((k_place_image<4, true>), dim3(static_cast<unsigned>(blocks)), dim3(kImgThreads), 0,
                                   rrint::stream(c), A);
            });
        } else {
            const void *kfn = A.channels == 4 ? reinterpret_cast<const void *>(&k_place_image<4, false>)
                                              : reinterpret_cast<const void *>(&k_place_image<3, false>);
            rc = blocks_of(kfn, lds_bytes, A.n_tiles, &blocks);
            if (rc) return rc;
            rc = launch_profiled(c, [&] {
                if (A.channels == 4)
                    hipLaunchKernelGGL((k_place_image<4, false>), dim3(static_cast<unsigned>(blocks)), dim3(kImgThreads),
                                       lds_bytes, rrint::stream(c), A);
                else
                    hipLaunchKernelGGL((k_place_image<3, false>), dim3(static_cast<unsigned>(blocks)), dim3(kImgThreads),
                                       lds_bytes, rrint::stream(c), A);
            });
        }
        if (rc) return rc;
    }
    return RR_OK;
}

int32_t rr_hip_record_icc_clut(rr_hip_ctx *c, const uint8_t *cmyk, uint8_t *rgb, const uint8_t *clut, size_t clut_len,
                               int64_t n_pixels) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (n_pixels < 0 || (n_pixels > 0 && (!cmyk || !rgb)) || (reinterpret_cast<uintptr_t>(cmyk) & 3u))
        return rrint::fail(c, RR_E_INVARIANT, "icc_clut: bad buffers (cmyk must be 4-byte aligned)");
    // grid_n_from_clut_len (crates/gpu/src/backend/params.rs:127-147): the
    // table must be exactly grid_n^4 RGB nodes with 2 <= grid_n <= 255
    int32_t grid_n = 0;
    if (clut) {
        if (clut_len % 3 == 0) {
            const size_t nodes = clut_len / 3;
            for (size_t g = 2; g <= 255; ++g) {
                const size_t p4 = g * g * g * g;
                if (p4 == nodes) grid_n = static_cast<int32_t>(g);
                if (p4 >= nodes) break;
            }
        }
        if (grid_n == 0) return rrint::fail(c, RR_E_INVARIANT, "icc_clut: table length is not grid_n^4 * 3 (2 <= grid_n <= 255)");
    }
    if (n_pixels == 0) return RR_OK;
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    if (!clut)
        return launch_profiled(c, [&] {
            hipLaunchKernelGGL(k_cmyk_matrix, dim3(grid_for(n_pixels)), dim3(256), 0, rrint::stream(c), cmyk, rgb, n_pixels);
        });
    const float scale = static_cast<float>(grid_n - 1) / 255.0f;  // cmyk.rs: g1 / 255.0 in f32
    return launch_profiled(c, [&] {
        hipLaunchKernelGGL(k_icc_clut, dim3(grid_for(n_pixels)), dim3(256), 0, rrint::stream(c), cmyk, rgb, clut, grid_n,
                           scale, n_pixels);
    });
}

int32_t rr_hip_record_composite(rr_hip_ctx *c, const uint8_t *src, uint8_t *dst, int64_t n_pixels) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (n_pixels < 0 || (n_pixels > 0 && (!src || !dst)) || ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 3u))
        return rrint::fail(c, RR_E_INVARIANT, "composite: bad buffers (4-byte aligned RGBA required)");
    if (n_pixels == 0) return RR_OK;
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    return launch_profiled(c, [&] {
        hipLaunchKernelGGL(k_composite_rgba8, dim3(grid_for(n_pixels)), dim3(256), 0, rrint::stream(c), src, dst, n_pixels);
    });
}

int32_t rr_hip_record_rgb_to_gray(rr_hip_ctx *c, const uint8_t *rgb, int64_t rgb_stride, uint8_t *gray,
                                  int64_t gray_stride, int32_t width, int32_t height) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (width < 0 || height < 0 || height > 65535 || rgb_stride < 3ll * width || gray_stride < width ||
        (width > 0 && height > 0 && (!rgb || !gray)))
        return rrint::fail(c, RR_E_INVARIANT, "rgb_to_gray: bad bitmap geometry");
    if (width == 0 || height == 0) return RR_OK;
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    const int32_t aligned = ((reinterpret_cast<uintptr_t>(rgb) | reinterpret_cast<uintptr_t>(gray) |
                              static_cast<uintptr_t>(rgb_stride) | static_cast<uintptr_t>(gray_stride)) & 3u) == 0;
    const dim3 grid(static_cast<unsigned>((width + 1023) / 1024), static_cast<unsigned>(height));
    return launch_profiled(c, [&] {
        hipLaunchKernelGGL(k_rgb_to_gray, grid, dim3(256), 0, rrint::stream(c), rgb, rgb_stride, gray, gray_stride,
                           width, height, aligned);
    });
}

int32_t rr_hip_record_apply_soft_mask(rr_hip_ctx *c, uint8_t *pixels, const uint8_t *mask, int64_t n_pixels) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (n_pixels < 0 || (n_pixels > 0 && (!pixels || !mask)))
        return rrint::fail(c, RR_E_INVARIANT, "soft_mask: bad buffers");
    if (n_pixels == 0) return RR_OK;
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    return launch_profiled(c, [&] {
        hipLaunchKernelGGL(k_apply_soft_mask, dim3(grid_for(n_pixels)), dim3(256), 0, rrint::stream(c), pixels, mask,
                           n_pixels);
    });
}

}  // extern "C"
