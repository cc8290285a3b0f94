/*
 * rr_hip.h -- C-ABI of the MI355X (gfx950) HIP backend for the rasterrocket
 * GPU raster hot path: anti-aliased path fill + span compositing on device.
 *
 * Every entry point returns int32_t: 0 on success, a negative RR_E_* code on
 * failure (the classes of BackendError, crates/gpu/src/backend/mod.rs:35-72);
 * rr_hip_last_error(ctx) gives the message.  Callers turn any error into the
 * reference's CPU fallback, as crates/pdf_interp/src/renderer/page/vk_ops.rs:
 * 62-70 does for the Vulkan backend.  Buffers are owned by the caller; a
 * context serialises its work on one HIP stream.
 */
#ifndef RR_HIP_H
#define RR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error classes (BackendError variants, backend/mod.rs:35-72). */
#define RR_OK 0
#define RR_E_ZERO_SIZE_ALLOC (-1) /* BackendError::ZeroSizeAlloc */
#define RR_E_INVARIANT (-2)       /* BackendError::InvariantViolation */
#define RR_E_DEVICE (-3)          /* BackendError::Other: a HIP runtime failure */
#define RR_E_NO_DEVICE (-4)       /* backend unavailable (init failure) */
#define RR_E_CAPACITY (-5)        /* device scratch pool exhausted for this batch */

/* AA row mode. DOCUMENTED: one renderAALine per device row, the Splash
 * semantics the reference documents (crates/raster/src/scanner/mod.rs:380-384).
 * LITERAL: crates/raster/src/fill/mod.rs:176-203 as written (output row y
 * receives device row 4y+3). */
#define RR_MODE_DOCUMENTED 0
#define RR_MODE_LITERAL 1
/* Canvas flavour. GRAY8: Bitmap<Gray8>, exact div255 (pipe/aa.rs:192-208).
 * RGB8_FAST: Bitmap<Rgb8> on a grey page, (v+255)>>8 (simd/composite.rs:56-114),
 * stored one byte per pixel because r=g=b stays invariant.
 * RGB8: Bitmap<Rgb8> as raster_pdf renders it (crates/pdf_interp/src/renderer/
 * page/mod.rs:479-480): 3 bytes a pixel (r, g, b), every path with its own
 * colour (rr_fill_pages.colors), composited per channel with the same
 * formulas; rows of `stride` bytes, stride a multiple of 4 and at least
 * 3 * (width rounded up to 4).  A page's optional Gray8 output receives
 * rgb_to_gray of the finished page (crates/pdf_raster/src/render.rs:1306-1324). */
#define RR_FLAVOUR_GRAY8 0
#define RR_FLAVOUR_RGB8_FAST 1
#define RR_FLAVOUR_RGB8 2
/* Execution shape of record_fill_paths. */
#define RR_LAUNCH_PLAIN 0   /* one path per kernel launch; launches of paths whose drawn rows \
                               meet run in path order, the others may run concurrently */
#define RR_LAUNCH_BATCHED 1 /* one persistent launch for all paths of the batch */

/* Per-path fill attributes (device array).  Mirrors the arguments of
 * raster::fill / eo_fill (crates/raster/src/fill/mod.rs:55,83) for a solid
 * source: fill rule, grey value, PipeState::a_input, and the two PipeState
 * fields that send a fill to the general pipe instead of the AA path
 * (render_span, crates/raster/src/pipe/mod.rs:157-199): blend mode Multiply
 * and a soft mask.  The general pipe composites with compute_a_src and the
 * no-alpha formulas of crates/raster/src/pipe/general.rs:270-344 (Multiply:
 * blend_multiply, pipe/blend.rs:24-26); the soft mask is the page's
 * (rr_page_desc.soft_mask), read at each pixel. */
#define RR_FILL_EVEN_ODD 1  /* even-odd rule (else non-zero winding) */
#define RR_FILL_MULTIPLY 2  /* BlendMode::Multiply (else Normal) */
#define RR_FILL_SOFT_MASK 4 /* PipeState::soft_mask = the page's soft mask */
typedef struct {
    uint8_t eo;      /* RR_FILL_* bits; 1 alone = even-odd, 0 = non-zero winding */
    uint8_t color;   /* solid grey source value (RR_FLAVOUR_RGB8: see colors) */
    uint8_t a_input; /* fill alpha, 255 = opaque */
    uint8_t route;   /* fill route chosen by the dispatcher (RR_ROUTE_*) */
} rr_path_attr;

/* Fill routes (the dispatcher's classes).  Every route yields the same pixels.
 * SPAN: the batched span pass takes the path one 16-row band per work item.
 * TILE16/32/64: one wave takes all of the path's bands (at most 1/2/4) in one
 *   work item; a path with more bands falls back to SPAN.
 * CPU: the path is best filled by the caller's CPU raster (the static rule's
 *   class for bbox area < 256 px, crates/gpu/src/lib.rs:112,120); handed to
 *   the library anyway, it is filled like SPAN.
 * AUTO: chosen in-kernel ahead of the fill: by the dispatcher MLP when one is
 *   loaded (rr_hip_set_dispatcher), else by the static rule (area < 256 px ->
 *   CPU, else the smallest TILE that holds the path's bands, else SPAN).
 * Routes are honoured by RR_LAUNCH_BATCHED.  RR_LAUNCH_PLAIN takes its shape
 * from the path's host_rows instead (the route lives in device memory): a path
 * of two or three 16-row bands runs as one workgroup holding all its bands
 * (the tile shape of a small path, a launch the device dispatches fastest),
 * any other one workgroup per band. */
#define RR_ROUTE_SPAN 0
#define RR_ROUTE_TILE16 1
#define RR_ROUTE_TILE32 2
#define RR_ROUTE_TILE64 3
#define RR_ROUTE_CPU 4
#define RR_ROUTE_AUTO 255

/* One page of paths to fill and composite, in order.  Replaces the per-path
 * AaFillParams / TileFillParams (crates/gpu/src/backend/params.rs:90-103,
 * 170-185) + CPU gpu_coverage_to_bitmap (renderer/page/gpu_ops.rs:118-220):
 * the coverage never leaves the device.
 *
 * edges: device, 4 doubles per edge = the (x0,y0,x1,y1) arguments of every
 *   XPath::add_segment call (crates/raster/src/xpath.rs:324) in order, device
 *   space, before aa_scale.  rr_xpath_edges() produces them.
 * edge_off: device, n_paths+1 int64 offsets into edges.
 * host_rows: host, 2 int32 per path = [first, last] device row touched by the
 *   path's edges (floor of min/max y); only used to size plain launches.
 * cov_hash: optional device array of n_paths uint64; each receives the sum of
 *   rr_cov_hash_term() over the pixels the path composited with coverage > 0.
 * cov_count: optional device array of 2*n_paths uint64; [2p] receives the
 *   number of pixels path p composited with coverage > 0, [2p+1] the sum over
 *   its drawn rows of the reference's draw_aa_line width x1-x0+1
 *   (crates/raster/src/fill/mod.rs:344-401), i.e. the pixel-composites the
 *   reference CPU path performs.
 */
typedef struct {
    const double *edges;
    const int64_t *edge_off;
    const rr_path_attr *attrs;
    const int32_t *host_rows;
    int64_t n_paths;
    int64_t n_edges;
    uint8_t *canvas; /* device, 1 byte per pixel, row pitch `stride` */
    int32_t width, height, stride;
    int32_t mode, flavour, launch;
    double clip[4]; /* x_min, y_min, x_max, y_max; page clip is {0,0,W-0.001,H-0.001} */
    uint64_t *cov_hash;
    uint64_t *cov_count;
    int32_t variant; /* band-fill shape, 0 <= variant < rr_hip_variant_count() */
    int32_t reserved;
    const uint8_t *colors;     /* RR_FLAVOUR_RGB8: device, 4 bytes a path {r, g, b, 0}, 4-byte aligned */
    const uint8_t *soft_mask;  /* device page-space soft mask, 1 byte a pixel, or NULL */
    int64_t soft_mask_stride;  /* bytes a row, a multiple of 4 (the mask 4-byte aligned) */
    uint8_t *gray;             /* RR_FLAVOUR_RGB8: optional device Gray8 output, or NULL */
    int64_t gray_stride;       /* bytes a row, a multiple of 4 >= width (the output 4-byte aligned) */
} rr_fill_batch;

/* One page of a multi-page batch: its canvas and the range of the batch's
 * paths that belong to it (pages tile [0, n_paths) in order).  With a gray
 * output, the batched launch writes it in the composite pass's epilogue
 * (every tile of the page, touched or not); the plain launch converts the
 * page after its last path.  wait_page returns RR_E_INVARIANT when a path
 * asked for the soft mask of a page that has none (that path is filled as if
 * the mask were 255). */
typedef struct {
    uint8_t *canvas; /* device, 1 byte per pixel (3 for RR_FLAVOUR_RGB8), row pitch `stride` */
    int32_t width, height, stride, reserved;
    int64_t path_begin, path_end;
    double clip[4]; /* page clip is {0,0,W-0.001,H-0.001} (crates/raster/src/state.rs:108-121) */
    const uint8_t *soft_mask; /* as in rr_fill_batch; required by paths with RR_FILL_SOFT_MASK */
    int64_t soft_mask_stride;
    uint8_t *gray;            /* as in rr_fill_batch (RR_FLAVOUR_RGB8 only) */
    int64_t gray_stride;
} rr_page_desc;

/* Many pages of paths in one recording.  edges/edge_off/attrs/cov_* are as in
 * rr_fill_batch with path indices global to the batch; pages is a host array.
 * With launch = RR_LAUNCH_BATCHED every page goes through one edge-setup pass
 * and one persistent fill+composite launch; with RR_LAUNCH_PLAIN each path is
 * its own launch (host_rows required), spread over up to 20 HIP streams (two
 * fewer than GPU_MAX_HW_QUEUES when that is 8 or more; the variable must be set
 * before the HIP runtime starts) and ordered by the rows the paths draw. */
typedef struct {
    const double *edges;
    const int64_t *edge_off;
    const rr_path_attr *attrs;
    const int32_t *host_rows;
    int64_t n_paths;
    int64_t n_edges;
    const rr_page_desc *pages;
    int32_t n_pages;
    int32_t mode, flavour, launch;
    uint64_t *cov_hash;
    uint64_t *cov_count;
    int32_t variant;
    int32_t reserved;
    const uint8_t *colors; /* as in rr_fill_batch, indexed by the batch's path */
} rr_fill_pages;



/* ---- context (GpuCtx::init, crates/gpu/src/lib.rs:239; CudaBackend, backend/cuda/mod.rs:27) ---- */
typedef struct rr_hip_ctx rr_hip_ctx;
int32_t rr_hip_create(int32_t device, rr_hip_ctx **out);
int32_t rr_hip_destroy(rr_hip_ctx *ctx);
const char *rr_hip_last_error(rr_hip_ctx *ctx);

/* ---- buffers (GpuBackend::alloc_device / alloc_device_zeroed / free_device /
 *      device_buffer_len, backend/mod.rs:241-282) ---- */
int32_t rr_hip_alloc(rr_hip_ctx *ctx, size_t size, int32_t zeroed, void **buf);
int32_t rr_hip_free(rr_hip_ctx *ctx, void *buf);
/* GpuBackend::device_buffer_len (backend/mod.rs:267): the size a buffer was
 * allocated with by rr_hip_alloc (RR_E_INVARIANT for any other pointer). */
int32_t rr_hip_buffer_len(rr_hip_ctx *ctx, const void *buf, size_t *len);
/* GpuBackend::alloc_host_pinned / free_host_pinned (backend/mod.rs:292-299):
 * page-locked host memory for DMA transfers; size 0 is RR_E_ZERO_SIZE_ALLOC. */
int32_t rr_hip_alloc_host_pinned(rr_hip_ctx *ctx, size_t size, void **buf);
int32_t rr_hip_free_host_pinned(rr_hip_ctx *ctx, void *buf);
/* GpuBackend::detect_vram_budget (backend/mod.rs:502; VramBudget :157-166):
 * total device memory and the usable share, 3/4 of what the driver reports
 * free (as the CUDA backend computes it, backend/cuda/mod.rs:211-220). */
int32_t rr_hip_vram_budget(rr_hip_ctx *ctx, uint64_t *total_bytes, uint64_t *usable_bytes);

/* ---- transfer (backend/mod.rs:422-496) ----
 * upload / download: synchronous copies.  The _async forms enqueue the copy on
 * the context stream, behind everything recorded before it, and return a
 * fence (upload_async / download_async; the destination of a download holds
 * the bytes once the fence is waited).  submit_transfer returns a fence after
 * every transfer issued so far; wait_transfer blocks on any fence of the
 * context (GpuBackend::wait_transfer / wait_download).  A fence is 1 + the
 * number of fences handed out before it; the last 16 are tracked. */
int32_t rr_hip_upload(rr_hip_ctx *ctx, void *dst, const void *src, size_t n);
int32_t rr_hip_download(rr_hip_ctx *ctx, void *dst, const void *src, size_t n);
int32_t rr_hip_upload_async(rr_hip_ctx *ctx, void *dst, const void *src, size_t n, uint64_t *fence);
int32_t rr_hip_download_async(rr_hip_ctx *ctx, void *dst, const void *src, size_t n, uint64_t *fence);
int32_t rr_hip_submit_transfer(rr_hip_ctx *ctx, uint64_t *fence);
int32_t rr_hip_wait_transfer(rr_hip_ctx *ctx, uint64_t fence);
int32_t rr_hip_memset(rr_hip_ctx *ctx, void *dst, int32_t value, size_t n);
int32_t rr_hip_copy(rr_hip_ctx *ctx, void *dst, const void *src, size_t n); /* device to device */

/* ---- page recording (begin_page / record_* / submit_page / wait_page,
 *      backend/mod.rs:306-405) ---- */
int32_t rr_hip_begin_page(rr_hip_ctx *ctx);
int32_t rr_hip_record_fill_paths(rr_hip_ctx *ctx, const rr_fill_batch *batch);
int32_t rr_hip_record_fill_pages(rr_hip_ctx *ctx, const rr_fill_pages *batch);
/* submit_page returns the page's fence; wait_page blocks on it and returns
 * that page's status (RR_E_CAPACITY when a device workspace overflowed and the
 * page output is incomplete), which later pages cannot overwrite while the
 * fence is among the last 16 handed out. */
int32_t rr_hip_submit_page(rr_hip_ctx *ctx, uint64_t *fence);
int32_t rr_hip_wait_page(rr_hip_ctx *ctx, uint64_t fence);
/* Fault injection (no counterpart in the reference; for testing the callers'
 * CPU fallback, crates/pdf_interp/src/renderer/page/vk_ops.rs:62-70): ORs
 * `bits` (nonzero; 1 = overflow-entry pool exhausted) into the open page's
 * error word, ordered after the work recorded so far, so that wait_page
 * reports the page as incomplete exactly as a real overflow would. */
int32_t rr_hip_inject_error(rr_hip_ctx *ctx, int32_t bits);
/* Timing of the last waited page, from HIP events on the context stream:
 * span_ms from the first to the last recorded kernel; with profiling enabled
 * (rr_hip_set_profiling) fill_ms is the summed duration of the fill kernels
 * alone, each bracketed by its own event pair (RR_LAUNCH_PLAIN then runs its
 * launches one after another instead of concurrently, and RR_LAUNCH_BATCHED
 * builds its band lists in line instead of beside the span pass); launches
 * counts the fill-kernel launches (the edge-setup pass excluded). */
int32_t rr_hip_set_profiling(rr_hip_ctx *ctx, int32_t enable);
/* Per-fill timing trace (no counterpart in the reference; it feeds the
 * autotuner and the learned dispatcher).  While set, every fill of one path
 * over one 16-row band appends a record of four int64 {path index,
 * (page << 32) | first output row, start, end} (device wall-clock ticks,
 * 100 MHz) to the caller's device buffer, up to `capacity` records; capacity
 * 0 turns tracing off.  rr_hip_trace_count reports the records written since
 * the last set (it may exceed capacity: the rest were dropped). */
int32_t rr_hip_set_trace(rr_hip_ctx *ctx, void *records, int64_t capacity);
int32_t rr_hip_trace_count(rr_hip_ctx *ctx, int64_t *count);

/* Band-fill shapes the library is built with (the autotuner's search space;
 * no counterpart in the reference).  Every variant gives identical pixels;
 * they differ in the LDS entry slots per AA row and the composite chunk
 * width, hence in occupancy.  variant_count() is the number of variants. */
int32_t rr_hip_variant_count(void);
int32_t rr_hip_variant_info(int32_t variant, int32_t *lds_entries, int32_t *chunk_px);
/* Autotuner (no counterpart in the reference): records `batch` once per
 * variant plus `reps` timed repetitions (HIP events around the recorded
 * work), writes the mean milliseconds of each variant to ms[] (variant_count
 * entries, may be NULL) and the fastest variant to *best.  The batch's
 * canvases are overwritten: reset them before the real run. */
int32_t rr_hip_autotune_fill(rr_hip_ctx *ctx, const rr_fill_pages *batch, int32_t reps, int32_t *best,
                             double *ms);
int32_t rr_hip_last_page_stats(rr_hip_ctx *ctx, double *span_ms, double *fill_ms, int64_t *launches);

/* ---- host path ingestion (XPath::new, crates/raster/src/xpath.rs:162-260,
 *      with flatten_curve, path/flatten.rs:109-215) ----
 * Turns a PathBuilder point list (pts xy pairs + PathFlags bytes,
 * path/mod.rs:70-79) into the add_segment edge list the device consumes.
 * Writes at most cap edges; *n_out receives the full count. */
int32_t rr_xpath_edges(const double *xy, const uint8_t *flags, int64_t n_pts,
                       const double matrix[6], double flatness, double *out_edges, int64_t cap,
                       int64_t *n_out);

/* ---- device path ingestion (no counterpart in the reference, which runs
 *      XPath::new on the CPU; the same walk and flatten_curve on the GPU) ----
 * Every path p is points [pt_off[p], pt_off[p+1]) of xy (2 doubles each, user
 * space) and flags (PathFlags bytes), transformed by matrix[6p..6p+6) (NULL:
 * identity) as XPath::new does (crates/raster/src/xpath.rs:162-260,
 * 421-426; close_subpaths = true, the fill case) and flattened with
 * flatten_curve (crates/raster/src/path/flatten.rs:109-215).  Writes path p's
 * edges, bit-identical to rr_xpath_edges, to edges[4*edge_off[p] ..
 * 4*edge_off[p+1]) -- the layout rr_fill_pages takes.  All pointers are
 * device memory.  Synchronous; not inside an open page.  *n_edges receives
 * the total; when it exceeds edge_cap nothing is written to edges and the
 * call returns RR_E_CAPACITY (edge_off is written either way).  pt_off must
 * rise from 0 to n_pts (RR_E_INVARIANT otherwise). */
typedef struct rr_xpath_batch {
    const double *xy;
    const uint8_t *flags;
    const int64_t *pt_off;
    const double *matrix;
    int64_t n_paths, n_pts;
    double flatness;
    double *edges;
    int64_t edge_cap;
    int64_t *edge_off;
} rr_xpath_batch;
int32_t rr_hip_xpath_edges(rr_hip_ctx *ctx, const rr_xpath_batch *batch, int64_t *n_edges);
/* Kernel milliseconds of the last rr_hip_xpath_edges call (walk, count, emit;
 * HIP events, filled while per-launch profiling is on, else 0); no
 * counterpart in the reference. */
int32_t rr_hip_last_xpath_stats(rr_hip_ctx *ctx, double ms[3]);

/* Coverage-hash term of one composited pixel (index y*width+x, shape 1..255). */
uint64_t rr_cov_hash_term(int64_t pixel_index, uint8_t shape);

/* ---- synthetic workloads (seeded; SplitMix64) ----
 * config 1: 500 random polygons on one Letter@300 page, half NZ / half EO.
 * config 2: 6000 synthetic glyph outlines on one Letter@300 page.
 * config 3: the mixed 64-page batch (Letter/A3/A4@600/thumbnail).
 * config 5: a stream of max_pages (32 if 0) glyph pages of the config-2
 *   distribution, one generator across the stream.
 * A workload holds, per page, the PathBuilder point lists and the edge list. */
typedef struct rr_workload rr_workload;
int32_t rr_workload_create(int32_t config, uint64_t seed, int32_t max_pages, rr_workload **out);
void rr_workload_destroy(rr_workload *w);
int32_t rr_workload_n_pages(const rr_workload *w);
/* width, height, n_paths, n_edges, n_points of page p. */
int32_t rr_workload_page_info(const rr_workload *w, int32_t p, int64_t info[5]);
/* Borrowed host pointers valid until rr_workload_destroy. */
int32_t rr_workload_page_data(const rr_workload *w, int32_t p, const double **edges,
                              const int64_t **edge_off, const rr_path_attr **attrs,
                              const int32_t **host_rows, const double **pts,
                              const uint8_t **pt_flags, const int64_t **pt_off);

/* ---- images + colour (image placement, soft mask, Normal/Multiply, CLUT) ---- */

/* One image XObject to place.  pixels: `channels` interleaved bytes per pixel
 * (3 = RGB, 4 = CMYK), row-major; smask: one byte per image pixel (or NULL =
 * opaque), sampled at the same image pixel as the colour
 * (crates/pdf_interp/src/renderer/page/mod.rs:1917-1920).  ctm: the PDF CTM
 * mapping the unit square to PDF user space on a page of the canvas height
 * (y up), as blit_image receives it (page/mod.rs:1600-1676).  blend: 0 Normal,
 * 1 Multiply (blend_multiply, crates/raster/src/pipe/blend.rs:24-26, then
 * blend_u8 under the mask).  page: index into the batch's page table. */
typedef struct {
    const uint8_t *pixels;
    const uint8_t *smask;
    int32_t width, height, channels, blend;
    double ctm[6];
    int32_t page, reserved;
} rr_image_desc;

/* Images applied in order onto the pages' canvases (device pointers in the
 * descriptors, host descriptor arrays).  grid4: CMYK->grey u16 grid with
 * grid4_n nodes per axis, index ((k*n + c)*n + m)*n + y; grid3: RGB->grey u16
 * grid, index (r*n + g)*n + b.  Sampling follows blit_image's CPU path: its
 * axis-aligned Q32 stepping when |b|*w < 0.5 and |c|*h < 0.5
 * (crates/pdf_interp/src/renderer/page/mod.rs:1767-1898), its general affine
 * path otherwise (:1899-1974).  Every descriptor is checked before anything is
 * launched: a rejected batch leaves the canvases untouched. */
typedef struct {
    const rr_image_desc *images;
    int32_t n_images;
    const rr_page_desc *pages;
    int32_t n_pages;
    const uint16_t *grid4;
    int32_t grid4_n;
    const uint16_t *grid3;
    int32_t grid3_n;
} rr_image_batch;
int32_t rr_hip_record_place_images(rr_hip_ctx *ctx, const rr_image_batch *batch);

/* Reference GPU ops over device buffers (GpuBackend::record_icc_clut /
 * record_composite / record_apply_soft_mask, crates/gpu/src/backend/mod.rs:
 * 321-336, params.rs:109-205), bit-exact to their CPU twins:
 * icc_cmyk_to_rgb_cpu (crates/gpu/src/cmyk.rs:523-634: clut == NULL selects the
 * reflectance matrix, color/src/convert.rs:131-161), composite_rgba8_cpu and
 * apply_soft_mask_cpu (crates/gpu/src/composite.rs:4-52).  The CLUT's grid_n
 * is recovered from clut_len (bytes) as grid_n_from_clut_len does
 * (params.rs:127-147): clut_len must be grid_n^4 * 3 with 2 <= grid_n <= 255,
 * else RR_E_INVARIANT. */
int32_t rr_hip_record_icc_clut(rr_hip_ctx *ctx, const uint8_t *cmyk, uint8_t *rgb, const uint8_t *clut,
                               size_t clut_len, int64_t n_pixels);
int32_t rr_hip_record_composite(rr_hip_ctx *ctx, const uint8_t *src, uint8_t *dst, int64_t n_pixels);
int32_t rr_hip_record_apply_soft_mask(rr_hip_ctx *ctx, uint8_t *pixels, const uint8_t *mask, int64_t n_pixels);

/* Page epilogue rgb_to_gray (crates/pdf_raster/src/render.rs:1306-1324, the
 * BT.709 conversion raster_pdf applies to its Rgb8 page): gray[y][x] =
 * (2126 r + 7152 g + 722 b + 5000) / 10000 of the Rgb8 bitmap (rows of
 * rgb_stride bytes, width * 3 used) into a Gray8 bitmap (rows of gray_stride
 * bytes), height <= 65535.  Identity on grey pixels: a grey canvas rendered
 * with RR_FLAVOUR_RGB8_FAST already is raster_pdf's output. */
int32_t rr_hip_record_rgb_to_gray(rr_hip_ctx *ctx, const uint8_t *rgb, int64_t rgb_stride, uint8_t *gray,
                                  int64_t gray_stride, int32_t width, int32_t height);

/* HIP-graph capture of a batched recording (no reference counterpart: the
 * batched pattern as one replayable graph).  graph_create runs an eager
 * sizing pass (edge setup, span pass; the canvases are not touched), sizes the
 * workspace, then captures the whole device sequence of rr_hip_record_fill_pages
 * for *batch (launch must be RR_LAUNCH_BATCHED; the device buffers and page
 * table it names are baked in).  graph_launch replays it inside
 * begin_page/submit_page; it fails with RR_E_INVARIANT once any recording on
 * the context has reallocated the workspace (capture again).  The caller
 * clears canvases / coverage buffers between replays as for eager records. */
typedef struct rr_hip_graph rr_hip_graph;
int32_t rr_hip_graph_create(rr_hip_ctx *ctx, const rr_fill_pages *batch, rr_hip_graph **out);
int32_t rr_hip_graph_launch(rr_hip_ctx *ctx, rr_hip_graph *graph);
int32_t rr_hip_graph_destroy(rr_hip_ctx *ctx, rr_hip_graph *graph);

/* ---- learned dispatcher (no reference counterpart) ----
 * MLP: 12 features -> 64 -> 64 -> 5 (ReLU), FP32; the outputs predict log2 of
 * the span-pass microseconds of each class (SPAN, TILE16, TILE32, TILE64,
 * CPU); AUTO takes the smallest prediction among the classes valid for the
 * path.  Parameters: RR_MLP_PARAMS floats = feature mean[12], feature
 * scale[12], W1[64][12], b1[64], W2[64][64], b2[64], W3[5][64], b3[5].
 * Features of path p (device pixels; W = column range, H = drawn rows,
 * E = edges): log2(1+E), log2(1+W), log2(1+H), log2(1+W*H),
 * log2(1+E/(4H)), bands, even-odd, log2((W+1)/(H+1)), |shoelace area|/(W*H)
 * clamped to 1, log2(1+AA rows per edge), a_input/255, log2(1+W*H/page area). */
#define RR_MLP_FEATURES 12
#define RR_MLP_CLASSES 5
#define RR_MLP_PARAMS 5341
/* Load (params != NULL, n == RR_MLP_PARAMS) or clear (params == NULL) the MLP. */
int32_t rr_hip_set_dispatcher(rr_hip_ctx *ctx, const float *params, int32_t n);
/* Features and chosen classes of the last batched recording: out_features
 * (host, n_paths*12 floats, may be NULL), out_routes (host, n_paths bytes,
 * may be NULL). */
int32_t rr_hip_last_dispatch(rr_hip_ctx *ctx, float *out_features, uint8_t *out_routes, int64_t n_paths);
/* MLP forward pass on the device: features (device, n*12) -> out (device,
 * n*5) with the loaded parameters. */
int32_t rr_hip_mlp_forward(rr_hip_ctx *ctx, const float *features, float *out, int64_t n);
/* Trains the dispatcher MLP on the device (one workgroup, Adam, masked
 * squared error): features (device, n*12), targets (device, n*5 log2
 * microseconds, NaN where a class is not valid), `steps` minibatches of
 * `batch` (<= 64) samples drawn with `seed`, learning rate lr.  params_out
 * (host, RR_MLP_PARAMS floats) receives the trained parameters (feature
 * standardisation from the data, He-initialised weights; a class with no
 * valid target keeps output bias 64, i.e. is never chosen); loss_out (host,
 * may be NULL) the mean masked loss of the last 100 steps. */
int32_t rr_hip_train_dispatcher(rr_hip_ctx *ctx, const float *features, const float *targets, int64_t n,
                                int32_t steps, int32_t batch, float lr, uint64_t seed, float *params_out,
                                float *loss_out);

/* Statistics of the last batched recording (no reference counterpart; read
 * after wait_page).  ms: {main span pass (k_spans), the wide span passes
 * (deferred bands, then the widened paths' skipped bands), composite pass}
 * kernel milliseconds when per-launch profiling was on for that page, else 0.  counts: {(path, band)
 * span items, intervals stored, (page, band, tile) composite items,
 * (path, tile) composites, tiles holding at least one path, span-pass reruns
 * after growing the interval pool or the deferred queue, (path, band) items
 * skipped as occluded, (path, band) items deferred to the wide span pass,
 * paths with at least one band built, the edge records of those paths,
 * composite tiles started from an opaque cover's colour (the canvas tile is
 * not read), intervals the composite pass read}. */
int32_t rr_hip_last_batch_stats(rr_hip_ctx *ctx, double ms[3], int64_t counts[12]);
/* Diagnostics of the last batched recording (no reference counterpart):
 * copies up to cap elements of one internal table to host memory and sets
 * *count to its length.  table 0: per composite tile item, the highest opaque
 * path covering it fully (int32, -1 none); 1: tile-bin offsets (int64, 4 per
 * item + 1); 2: bin entries (int32 x 4 {span item, path, full-cover columns,
 * attribute bytes}); 3: span heads (16 bytes {int64 offset, int16 fx0, fx1,
 * cx0, cx1}); 4: first span item of each path (int64, n_paths + 1); 5: the
 * 64-byte path records of the span pass; 6: per span item, the interval
 * count of each of its 64 AA rows (uint16 x 64); 7: the interval pool (int32
 * x 2 [s, e) in AA columns, a head's offset indexes it). */
int32_t rr_hip_debug_batch_table(rr_hip_ctx *ctx, int32_t table, void *dst, int64_t cap, int64_t *count);

/* In-place exclusive prefix sum of len_elems uint32 (wrapping), replacing the
 * three record_scan phases of the reference (GpuBackend::record_scan,
 * crates/gpu/src/backend/mod.rs:338-353; ScanParams, params.rs:207-367;
 * crates/gpu/kernels/blelloch_scan.cu): one call records all three.
 * block_sums: device scratch of ceil(len_elems / 1024) uint32.  len_elems == 0
 * is an invariant violation, as in ScanParams::validate. */
int32_t rr_hip_record_scan(rr_hip_ctx *ctx, uint32_t *data, uint32_t *block_sums, uint32_t len_elems);
/* rr_hip_record_scan without a block_sums argument, allocating nothing (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid it).  The block sums live in a 64 KiB device array inside the
 * library, one per GPU, held until the process ends, so len_elems is at most
 * 16777216 (RR_E_INVARIANT above that).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (several contexts or streams)
 * race; calls recorded on one context are ordered by its stream.  No part of
 * the array needs clearing between calls. */
int32_t rr_hip_record_scan_static(rr_hip_ctx *ctx, uint32_t *data, uint32_t len_elems);

/* Seeded image workload: n_images images, images_per_page per 2550x3300 page,
 * image i is CMYK + Normal when i is even, RGB + Multiply when odd.  info =
 * {n_images, n_pages, page_w, page_h, grid4_n, grid3_n}; rr_images_get fills a
 * descriptor with host pointers. */
typedef struct rr_image_set rr_image_set;
int32_t rr_images_create(uint64_t seed, int32_t n_images, int32_t images_per_page, rr_image_set **out);
void rr_images_destroy(rr_image_set *set);
int32_t rr_images_info(const rr_image_set *set, int32_t info[6]);
int32_t rr_images_get(const rr_image_set *set, int32_t i, rr_image_desc *desc);
int32_t rr_images_grids(const rr_image_set *set, const uint16_t **grid4, const uint16_t **grid3);

#ifdef __cplusplus
}
#endif
#endif
