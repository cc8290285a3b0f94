/*
 * rr_oracle_img.c -- CPU oracle for the image / colour half of the hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see rr_oracle.h).  Restates, with file:line
 * citations, the reference's image placement (general affine path), blend_u8,
 * blend_multiply, the CPU twins of the reference GPU composite / soft-mask /
 * ICC kernels, and defines the integer grey CLUT that the image workload uses
 * (the reference has no u16 grid, no 3-D grid and no CLUT-to-grey; that part is
 * pinned by invariants in tests/test_oracle_img.py, not by the reference).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "rr_oracle.h"

/* blend_u8 (crates/pdf_interp/src/renderer/page/mod.rs:2146-2149) */
uint8_t rro_blend_u8(uint8_t src, uint8_t dst, uint32_t alpha) {
    uint32_t ia = 255u - alpha;
    return (uint8_t)((src * alpha + dst * ia) / 255u);
}

/* blend_multiply (crates/raster/src/pipe/blend.rs:24-26) */
uint8_t rro_blend_multiply(uint8_t src, uint8_t dst) { return rro_div255((uint32_t)src * dst); }

/* composite_rgba8_cpu (crates/gpu/src/composite.rs:4-37) */
void rro_composite_rgba8(const uint8_t *src, uint8_t *dst, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        const uint8_t *s = src + 4 * i;
        uint8_t *d = dst + 4 * i;
        uint32_t a_src = s[3];
        if (a_src == 0) continue;
        if (a_src == 255) {
            memcpy(d, s, 4);
            continue;
        }
        uint32_t a_dst = d[3];
        uint32_t inv = 255u - a_src;
        uint32_t a_out = a_src + (a_dst * inv + 127u) / 255u;
        if (a_out == 0) continue;
        for (int c = 0; c < 3; ++c) {
            uint32_t b = ((uint32_t)s[c] * a_src + (uint32_t)d[c] * a_dst * inv / 255u + a_out / 2u) / a_out;
            d[c] = (uint8_t)(b < 255u ? b : 255u);
        }
        d[3] = (uint8_t)a_out;
    }
}

/* apply_soft_mask_cpu (crates/gpu/src/composite.rs:39-52) */
void rro_apply_soft_mask(uint8_t *px, const uint8_t *mask, int64_t n) {
    for (int64_t i = 0; i < n; ++i) px[4 * i + 3] = (uint8_t)(((uint32_t)px[4 * i + 3] * mask[i] + 127u) / 255u);
}

/* cmyk_to_rgb_reflectance (crates/color/src/convert.rs:131-161) */
void rro_cmyk_matrix(const uint8_t *cmyk, uint8_t *rgb, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t inv_k = 255u - cmyk[4 * i + 3];
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = (uint8_t)(((255u - cmyk[4 * i + c]) * inv_k + 127u) / 255u);
    }
}

/* icc_cmyk_to_rgb_cpu, CLUT arm (crates/gpu/src/cmyk.rs:531-630): 4-D, u8 RGB
 * nodes, index (k*G^3 + c*G^2 + m*G + y)*3, f32 quadrilinear with
 * t.mul_add(b - a, a), lerp order Y, M, C per K face, then K, round half away. */
static float lerpf(float a, float b, float t) { return fmaf(t, b - a, a); }

void rro_icc_clut_ref(const uint8_t *cmyk, uint8_t *rgb, int64_t n, const uint8_t *table, int32_t g) {
    const int64_t g2 = (int64_t)g * g, g3 = g2 * g;
    const float g1 = (float)(g - 1);
    const float scale = g1 / 255.0f;
    for (int64_t i = 0; i < n; ++i) {
        const uint8_t *s = cmyk + 4 * i;
        float fc = (float)s[0] * scale, fm = (float)s[1] * scale, fy = (float)s[2] * scale, fk = (float)s[3] * scale;
        int ic0 = (int)fc, im0 = (int)fm, iy0 = (int)fy, ik0 = (int)fk;
        if (ic0 > g - 1) ic0 = g - 1;
        if (im0 > g - 1) im0 = g - 1;
        if (iy0 > g - 1) iy0 = g - 1;
        if (ik0 > g - 1) ik0 = g - 1;
        int ic1 = ic0 + 1 < g ? ic0 + 1 : g - 1, im1 = im0 + 1 < g ? im0 + 1 : g - 1;
        int iy1 = iy0 + 1 < g ? iy0 + 1 : g - 1, ik1 = ik0 + 1 < g ? ik0 + 1 : g - 1;
        float wc = fc - (float)ic0, wm = fm - (float)im0, wy = fy - (float)iy0, wk = fk - (float)ik0;
        for (int ch = 0; ch < 3; ++ch) {
#define NODE(ci, mi, yi, ki) ((float)table[((int64_t)(ki) * g3 + (int64_t)(ci) * g2 + (int64_t)(mi) * g + (yi)) * 3 + ch])
            float c0m0k0 = lerpf(NODE(ic0, im0, iy0, ik0), NODE(ic0, im0, iy1, ik0), wy);
            float c0m1k0 = lerpf(NODE(ic0, im1, iy0, ik0), NODE(ic0, im1, iy1, ik0), wy);
            float c1m0k0 = lerpf(NODE(ic1, im0, iy0, ik0), NODE(ic1, im0, iy1, ik0), wy);
            float c1m1k0 = lerpf(NODE(ic1, im1, iy0, ik0), NODE(ic1, im1, iy1, ik0), wy);
            float rk0 = lerpf(lerpf(c0m0k0, c0m1k0, wm), lerpf(c1m0k0, c1m1k0, wm), wc);
            float c0m0k1 = lerpf(NODE(ic0, im0, iy0, ik1), NODE(ic0, im0, iy1, ik1), wy);
            float c0m1k1 = lerpf(NODE(ic0, im1, iy0, ik1), NODE(ic0, im1, iy1, ik1), wy);
            float c1m0k1 = lerpf(NODE(ic1, im0, iy0, ik1), NODE(ic1, im0, iy1, ik1), wy);
            float c1m1k1 = lerpf(NODE(ic1, im1, iy0, ik1), NODE(ic1, im1, iy1, ik1), wy);
            float rk1 = lerpf(lerpf(c0m0k1, c0m1k1, wm), lerpf(c1m0k1, c1m1k1, wm), wc);
#undef NODE
            float out = lerpf(rk0, rk1, wk);
            if (out < 0.0f) out = 0.0f;
            if (out > 255.0f) out = 255.0f;
            rgb[3 * i + ch] = (uint8_t)roundf(out);
        }
    }
}

/* Grey CLUT of the image workload (own definition; no reference counterpart).
 * D = 3 (RGB, grid index (r*G + g)*G + b) or 4 (CMYK, index (k*G^3 + c*G^2 +
 * m*G + y), the reference's K-major order), u16 nodes.  Per channel
 * t = v*(G-1), i = t/255, f = t - 255*i (weights in 1/255 units, exact).
 * Simplex (Kuhn) interpolation: channels sorted by f descending (ties by
 * channel index), vertex k steps the first k sorted axes, weight
 * f_(k) - f_(k+1) with f_(0) = 255 and f_(D+1) = 0; the weights sum to 255.
 * Output = (sum + 32767) / 65535, i.e. round-half-up of node/257 scale. */
uint8_t rro_clut_gray(const uint8_t *px, int32_t dims, const uint16_t *grid, int32_t g) {
    int32_t i0[4] = {0, 0, 0, 0}, f[4] = {0, 0, 0, 0}, ord[4] = {0, 1, 2, 3};
    int64_t stride[4];
    if (dims == 4) {
        stride[0] = (int64_t)g * g;         /* c */
        stride[1] = g;                      /* m */
        stride[2] = 1;                      /* y */
        stride[3] = (int64_t)g * g * g;     /* k */
    } else {
        stride[0] = (int64_t)g * g;
        stride[1] = g;
        stride[2] = 1;
    }
    int64_t base = 0;
    for (int a = 0; a < dims; ++a) {
        int32_t t = (int32_t)px[a] * (g - 1);
        i0[a] = t / 255;
        f[a] = t - 255 * i0[a];
        base += i0[a] * stride[a];
        ord[a] = a;
    }
    /* stable insertion sort of axes by f descending */
    for (int a = 1; a < dims; ++a) {
        int v = ord[a], b = a;
        while (b > 0 && f[ord[b - 1]] < f[v]) {
            ord[b] = ord[b - 1];
            --b;
        }
        ord[b] = v;
    }
    uint32_t acc = (uint32_t)(255 - f[ord[0]]) * grid[base];
    int64_t cur = base;
    for (int k = 0; k < dims; ++k) {
        int a = ord[k];
        if (i0[a] + 1 < g) cur += stride[a];
        int32_t w = f[a] - (k + 1 < dims ? f[ord[k + 1]] : 0);
        acc += (uint32_t)w * grid[cur];
    }
    return (uint8_t)((acc + 32767u) / 65535u);
}

/* f64 -> i64 as Rust's `as i64` converts: truncation, saturating, NaN -> 0. */
static int64_t sat_f64_i64(double x) {
    if (isnan(x)) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}

/* f64::clamp(0.0, 1.0): NaN stays NaN. */
static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }

/* One canvas pixel from image pixel idx: soft mask 0 skips, 255 copies,
 * else blend_u8 (page/mod.rs:1839-1852,1941-1950); the colour is the grey
 * CLUT value, through blend_multiply first under Multiply.  Returns 1 when
 * the pixel was written. */
static int place_px(uint8_t *row, int64_t dx, const uint8_t *pixels, int64_t idx, int32_t channels,
                    const uint8_t *smask, int32_t blend, const uint16_t *grid, int32_t g) {
    const uint32_t alpha = smask ? smask[idx] : 255u;
    if (alpha == 0) return 0;
    uint8_t s = rro_clut_gray(pixels + idx * channels, channels, grid, g);
    const uint8_t dst = row[dx];
    if (blend == 1) s = rro_blend_multiply(s, dst);
    row[dx] = alpha == 255u ? s : rro_blend_u8(s, dst, alpha);
    return 1;
}

/* Places one image on a grey canvas as blit_image's CPU path does
 * (crates/pdf_interp/src/renderer/page/mod.rs:1607-1647 bbox, :1656-1676 inverse
 * CTM): the axis-aligned path when |b|*w < 0.5 and |c|*h < 0.5 (:1767-1898),
 * else the general affine path (:1899-1974).  The sampled pixel goes to grey
 * through rro_clut_gray and is composited with blend_u8 under the image-space
 * soft mask; blend = 1 replaces the source by blend_multiply(src, dst) first.
 * Returns the number of canvas pixels written. */
int64_t rro_place_image(uint8_t *canvas, int32_t width, int32_t height, int32_t stride, const uint8_t *pixels,
                        int32_t img_w, int32_t img_h, int32_t channels, const uint8_t *smask, const double ctm[6],
                        int32_t blend, const uint16_t *grid, int32_t g) {
    const double a = ctm[0], b = ctm[1], c = ctm[2], d = ctm[3], e = ctm[4], f = ctm[5];
    const double page_h = (double)height;
    double xs[4], ys[4];
    const double ux[4] = {0.0, 1.0, 0.0, 1.0}, uy[4] = {0.0, 0.0, 1.0, 1.0};
    for (int k = 0; k < 4; ++k) { /* ctm_transform (renderer/gstate.rs:64-69) */
        xs[k] = fma(a, ux[k], c * uy[k]) + e;
        ys[k] = fma(b, ux[k], d * uy[k]) + f;
        if (!isfinite(xs[k]) || !isfinite(ys[k])) return 0;
    }
    double xmn = fmin(fmin(xs[0], xs[1]), fmin(xs[2], xs[3]));
    double xmx = fmax(fmax(xs[0], xs[1]), fmax(xs[2], xs[3]));
    double ymn = fmin(fmin(page_h - ys[0], page_h - ys[1]), fmin(page_h - ys[2], page_h - ys[3]));
    double ymx = fmax(fmax(page_h - ys[0], page_h - ys[1]), fmax(page_h - ys[2], page_h - ys[3]));
    int64_t dx0 = sat_f64_i64(floor(xmn)), dx1 = sat_f64_i64(ceil(xmx)), dy0 = sat_f64_i64(floor(ymn)),
            dy1 = sat_f64_i64(ceil(ymx));
#define CL(v, hi) ((v) < 0 ? 0 : ((v) > (hi) ? (hi) : (v)))
    int64_t bx0 = CL(dx0, width), bx1 = CL(dx1, width), by0 = CL(dy0, height), by1 = CL(dy1, height);
#undef CL
    if (bx0 >= bx1 || by0 >= by1) return 0;
    const double det = fma(a, d, -(b * c));
    if (fabs(det) < 1e-12) return 0;
    const double inv_det = 1.0 / det;
    const double iw = (double)img_w, ih = (double)img_h;
    int64_t written = 0;
    if (fabs(b) * iw < 0.5 && fabs(c) * ih < 0.5) {
        /* axis-aligned path (page/mod.rs:1767-1898): per row v and iy in f64,
         * per column a Q32 fixed-point image x stepped by img_w / a with
         * wrapping adds, clamped to the image. */
        const double fp_scale = 4294967296.0;
        const int64_t ix_step = sat_f64_i64((iw / a) * fp_scale);
        const int64_t img_max_x = (int64_t)img_w - 1;
        for (int64_t dy = by0; dy < by1; ++dy) {
            const double dy_pdf = page_h - (double)dy;
            const double v = clamp01((dy_pdf - f) / d);
            const int64_t iy = (int64_t)fmin((1.0 - v) * ih, ih - 1.0);
            const int64_t row_base = iy * img_w;
            const double u0 = clamp01(((double)bx0 - e) / a);
            int64_t ix_fp = sat_f64_i64(u0 * iw * fp_scale);
            uint8_t *row = canvas + dy * stride;
            for (int64_t dx = bx0; dx < bx1; ++dx) {
                int64_t ix = ix_fp >> 32;
                ix = ix < 0 ? 0 : (ix > img_max_x ? img_max_x : ix);
                ix_fp = (int64_t)((uint64_t)ix_fp + (uint64_t)ix_step);
                written += place_px(row, dx, pixels, row_base + ix, channels, smask, blend, grid, g);
            }
        }
        return written;
    }
    for (int64_t dy = by0; dy < by1; ++dy) {
        const double dy_pdf = page_h - (double)dy;
        const double dy_rel = dy_pdf - f;
        const double u_row = (-c * dy_rel) * inv_det;
        const double v_row = (a * dy_rel) * inv_det;
        uint8_t *row = canvas + dy * stride;
        for (int64_t dx = bx0; dx < bx1; ++dx) {
            const double dx_rel = (double)dx - e;
            double u = fma(d * dx_rel, inv_det, u_row);
            double v = fma((-b) * dx_rel, inv_det, v_row);
            u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
            v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
            const int64_t ix = (int64_t)fmin(u * iw, iw - 1.0);
            const int64_t iy = (int64_t)fmin((1.0 - v) * ih, ih - 1.0);
            written += place_px(row, dx, pixels, iy * img_w + ix, channels, smask, blend, grid, g);
        }
    }
    return written;
}
