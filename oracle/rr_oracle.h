/*
 * rr_oracle.h -- CPU oracle for the rasterrocket GPU raster hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU rasteriser (crates/raster, crates/color of pthomasfournet/pdf-raster),
 * against which the HIP product path is compared, and the timed CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in the product library links or calls it.
 *
 * Every function cites the reference file:line it follows.  Parity status of
 * each function (which known-answer test, closed form or invariant pins it)
 * is listed next to its declaration.
 */
#ifndef RR_ORACLE_H
#define RR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PathFlags (crates/raster/src/path/mod.rs:70-79). */
enum { RRO_PATH_FIRST = 1, RRO_PATH_LAST = 2, RRO_PATH_CLOSED = 4, RRO_PATH_CURVE = 8 };
/* XPathFlags (crates/raster/src/xpath.rs:52-60). */
enum { RRO_SEG_HORIZ = 1, RRO_SEG_VERT = 2, RRO_SEG_FLIPPED = 4 };

/* AA row mode: documented = one renderAALine per device row (Splash);
 * literal = the loop of crates/raster/src/fill/mod.rs:176-203 read as written
 * (output row y receives device row 4y+3). */
enum { RRO_MODE_DOCUMENTED = 0, RRO_MODE_LITERAL = 1 };
/* Canvas flavour: gray8 exact div255 (pipe/aa.rs:192-208) or the Rgb8
 * opaque fast path (simd/composite.rs:56-114) on a grey page. */
enum { RRO_FLAVOUR_GRAY8 = 0, RRO_FLAVOUR_RGB8_FAST = 1, RRO_FLAVOUR_RGB8 = 2 };
/* Fill flags: even-odd rule, blend mode Multiply (PipeState::blend_mode), a
 * page-space soft mask (PipeState::soft_mask, sampled at each pixel).  Either
 * of the last two sends the paint to the general pipe (pipe/mod.rs:157-199). */
enum { RRO_FILL_EVEN_ODD = 1, RRO_FILL_MULTIPLY = 2, RRO_FILL_SOFT_MASK = 4 };

typedef struct {
    double x0, y0, x1, y1, dxdy;
    int32_t dxdy_fp;
    uint32_t flags;
} rro_seg;

typedef struct {
    int32_t x0, x1, count;
} rro_intersect;

/* ---- colour / rounding primitives (crates/color/src/convert.rs) ---- */
uint8_t rro_div255(uint32_t x);                 /* :50-57; pinned: exhaustive vs round(x/255) */
int32_t rro_splash_floor(double x);             /* :262-264; pinned: convert.rs:377-414 cases */
int32_t rro_splash_ceil(double x);              /* :284-286 */
int32_t rro_splash_round(double x);             /* :305-307 */
int32_t rro_fp_from_f64(double x);              /* scanner/mod.rs:51-66 */
extern const uint8_t RRO_AA_GAMMA[17];          /* fill/mod.rs:46-48; pinned: round((i/16)^1.5*255) */

/* ---- path builder (crates/raster/src/path/mod.rs:291-451) ---- */
typedef struct rro_path rro_path;
rro_path *rro_path_new(void);
void rro_path_free(rro_path *p);
int rro_path_move_to(rro_path *p, double x, double y);   /* 0 ok, 1 NoCurPt, 2 BogusPath */
int rro_path_line_to(rro_path *p, double x, double y);
int rro_path_curve_to(rro_path *p, double x1, double y1, double x2, double y2, double x3, double y3);
int rro_path_close(rro_path *p, int force);
int64_t rro_path_len(const rro_path *p);
void rro_path_get(const rro_path *p, double *xy, uint8_t *flags);
/* conic_to_cubic (crates/font/src/outline.rs:106-112). */
void rro_conic_to_cubic(const double p0[2], const double pc[2], const double p3[2],
                        double p1_out[2], double p2_out[2]);

/* ---- flattening (crates/raster/src/path/flatten.rs:109-215) ---- */
int64_t rro_flatten_curve(const double p[8], double flatness_sq, double *out_xy, int64_t cap);

/* ---- XPath (crates/raster/src/xpath.rs:162-260,324-402,421-426) ----
 * Emits the add_segment argument list (x0,y0,x1,y1 in device space, input
 * orientation) in order.  Returns the edge count (which may exceed cap; only
 * the first cap edges are written). */
int64_t rro_xpath_edges(const double *xy, const uint8_t *flags, int64_t n_pts,
                        const double m[6], double flatness, int close_subpaths,
                        double *out_edges, int64_t cap);
/* add_segment applied to an edge list (xpath.rs:324-402). */
void rro_add_segments(const double *edges, int64_t n, rro_seg *out);
/* aa_scale (xpath.rs:283-303). */
void rro_aa_scale(rro_seg *segs, int64_t n);

/* ---- scanner (crates/raster/src/scanner/mod.rs:131-248,393-642) ---- */
typedef struct rro_scanner rro_scanner;
rro_scanner *rro_scanner_new(const rro_seg *segs, int64_t n, int eo, int32_t clip_y_min,
                             int32_t clip_y_max);
void rro_scanner_free(rro_scanner *s);
void rro_scanner_bbox(const rro_scanner *s, int32_t out[4]); /* x_min, x_max, y_min, y_max */
int64_t rro_scanner_row(const rro_scanner *s, int32_t y, rro_intersect *out, int64_t cap);
int rro_scanner_test(const rro_scanner *s, int32_t x, int32_t y);
/* render_aa_line into a 4-row AaBuf of width aa_width bits (row pitch
 * ceil(aa_width/8) bytes, MSB-first).  Returns the number of rows on which it
 * had to stop a span at aa_width because the winding never returned to zero
 * (the reference loops forever there). */
int rro_scanner_render_aa_line(const rro_scanner *s, uint8_t *aabuf, int32_t aa_width,
                               int32_t *x0, int32_t *x1, int32_t y);
/* Scanner statistics used by the test-suite: number of rows, entries, rows
 * with more than 32 entries, and those of them with a tie on x0. */
void rro_scanner_stats(const rro_scanner *s, int64_t out[4]);

/* AaBuf::set_span (crates/raster/src/bitmap.rs:535-574). */
void rro_aabuf_set_span(uint8_t *row, int32_t width, int32_t x0, int32_t x1);
/* aa_coverage_span_scalar (crates/raster/src/simd/aa_coverage.rs:51-69). */
void rro_aa_coverage_span(const uint8_t *aabuf, int32_t row_bytes, int32_t x0, int32_t count,
                          uint8_t *shape);

/* ---- compositing (pipe/aa.rs:192-208; simd/composite.rs:56-114) ---- */
uint8_t rro_composite_gray8(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape);
uint8_t rro_composite_rgb8_fast(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape);
int32_t rro_rgb_to_gray(int32_t r, int32_t g, int32_t b);  /* render.rs:1306-1324 */
void rro_rgb_to_gray_page(const uint8_t *rgb, int32_t rgb_stride, uint8_t *gray, int32_t gray_stride, int32_t w,
                          int32_t h);
/* general pipe, no-alpha arm (pipe/general.rs:270-344; blend.rs:24-26);
 * pinned by general.rs:518-570 (compute_a_src cases), the no-alpha readings of
 * :572-611 and closed forms (tests/test_oracle_general.py) */
uint32_t rro_compute_a_src(uint32_t a_input, uint32_t soft, uint32_t shape, int has_soft_mask, int has_shape);
uint8_t rro_composite_general(uint8_t dst, uint8_t src, uint8_t a_input, uint8_t shape, uint8_t soft,
                              int has_soft_mask, int multiply);

/* ---- fill (crates/raster/src/fill/mod.rs:110-244,344-401) ----
 * Fills one path given its add_segment edge list (pre aa_scale) onto a 1-byte
 * canvas.  clip = {x_min, y_min, x_max, y_max} in device space (the page clip
 * is {0, 0, W-0.001, H-0.001}; state.rs:108-121).  When shape_out is not NULL
 * it receives the gamma-mapped coverage of every drawn pixel (W*H bytes,
 * untouched elsewhere).  cov_hash (if not NULL) is incremented by the
 * order-independent hash of every (pixel, shape>0) pair drawn.  Returns the
 * number of pixels composited with shape>0. */
int64_t rro_fill(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                 int64_t n_edges, int eo, uint8_t color, uint8_t a_input, int mode, int flavour,
                 const double clip[4], uint8_t *shape_out, uint64_t *cov_hash);

/* rro_fill with every paint option: flags RRO_FILL_*, colour r, g, b (a 1-byte
 * canvas uses color[0]), flavour RRO_FLAVOUR_RGB8 for a 3-byte Rgb8 canvas
 * (stride in bytes), soft_mask a W x H page-space mask (soft_stride bytes a
 * row) read when flags has RRO_FILL_SOFT_MASK. */
int64_t rro_fill_ex(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges, int64_t n_edges,
                    uint32_t flags, const uint8_t color[3], uint8_t a_input, int mode, int flavour,
                    const double clip[4], const uint8_t *soft_mask, int32_t soft_stride, uint8_t *shape_out,
                    uint64_t *cov_hash);
/* A page of rro_fill_ex paths (flags[p], colors[3p..3p+2], a_input[p]). */
int64_t rro_fill_page_ex(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                         const int64_t *edge_off, int64_t n_paths, const uint8_t *flags, const uint8_t *colors,
                         const uint8_t *a_input, int mode, int flavour, const uint8_t *soft_mask,
                         int32_t soft_stride, uint64_t *cov_hashes);

/* One page of paths, in order.  edge_off has n_paths+1 entries.  Returns the
 * total number of shape>0 pixels composited. */
int64_t rro_fill_page(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                      const int64_t *edge_off, int64_t n_paths, const uint8_t *eo,
                      const uint8_t *color, const uint8_t *a_input, int mode, int flavour,
                      uint64_t *cov_hashes);

/* rro_fill_page (Gray8, no hashes) with each path's wall time in ns[p]. */
int64_t rro_fill_page_timed(uint8_t *canvas, int32_t w, int32_t h, int32_t stride, const double *edges,
                            const int64_t *edge_off, int64_t n_paths, const uint8_t *eo, const uint8_t *color,
                            const uint8_t *a_input, int mode, int64_t *ns);

/* The reference dispatcher's bbox area (gpu_ops.rs:52-84), -1 = stays on the
 * CPU; pinned by hand-computed cases (tests/test_oracle_general.py). */
int64_t rro_ref_bbox_area(const double *edges, int64_t n, int32_t w, int32_t h);

/* Order-independent coverage hash term for one pixel. */
uint64_t rro_cov_hash_term(int64_t pixel_index, uint8_t shape);

/* ---- images and colour (rr_oracle_img.c) ---- */
uint8_t rro_blend_u8(uint8_t src, uint8_t dst, uint32_t alpha);      /* page/mod.rs:2146-2149; pinned :2366-2395 */
uint8_t rro_blend_multiply(uint8_t src, uint8_t dst);                /* pipe/blend.rs:24-26 */
void rro_composite_rgba8(const uint8_t *src, uint8_t *dst, int64_t n); /* gpu/src/composite.rs:4-37; pinned :58-83 */
void rro_apply_soft_mask(uint8_t *px, const uint8_t *mask, int64_t n); /* gpu/src/composite.rs:39-52; pinned :85-109 */
void rro_cmyk_matrix(const uint8_t *cmyk, uint8_t *rgb, int64_t n);  /* color/src/convert.rs:131-161 */
/* gpu/src/cmyk.rs:531-630; pinned by cmyk.rs:809-846 corners + closed forms */
void rro_icc_clut_ref(const uint8_t *cmyk, uint8_t *rgb, int64_t n, const uint8_t *table, int32_t g);
/* Grey CLUT of the image workload: own definition, pinned by invariants. */
uint8_t rro_clut_gray(const uint8_t *px, int32_t dims, const uint16_t *grid, int32_t g);
int64_t rro_place_image(uint8_t *canvas, int32_t width, int32_t height, int32_t stride, const uint8_t *pixels,
                        int32_t img_w, int32_t img_h, int32_t channels, const uint8_t *smask, const double ctm[6],
                        int32_t blend, const uint16_t *grid, int32_t g);

#ifdef __cplusplus
}
#endif
#endif
