// rr_internal.h -- the backend context and helpers shared by the translation
// units of the HIP backend (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "rr_device.h"
#include "rr_hip.h"

namespace rrdev {

// Device counters of a batched recording (unsigned long long each).
constexpr int kCounters = 16;
enum {
    kCntReserved0 = 0,     // (the work queues live in the sliced queue counters)
    kCntReserved1 = 1,
    kCntPoolTop = 2,       // interval pool bump pointer (records handed out)
    kCntStored = 3,        // intervals stored
    kCntPathTiles = 4,     // (path, tile) composites
    kCntTiles = 5,         // tiles holding at least one path
    kCntWidened = 6,       // paths widened by an unbalanced row
    kCntDeferred = 7,      // bands deferred to the wide pass
    kCntWideServed = 8,    // deferred bands served
    kCntFixupServed = 9,   // fixup items served
    kCntSkipped = 10,      // bands skipped as occluded
    kCntBuiltPaths = 11,   // statistics: paths with a built band, then their edges (two words)
    kCntFilledTiles = 13,  // composite tiles started from an opaque cover's colour (not read)
    kCntIvalsRead = 14,    // intervals the composite pass read
};

// Sliced work-queue counters of the span and composite passes (8 slices
// each, one 256-byte line per slice).
constexpr int kQueueWords = 2 * 8 * 32;

constexpr int kNumVariants = 16;  // instantiated fill shapes (the autotuner's search space)
// K = LDS entry slots per AA row; CH = pixels per composite sweep (plain) /
// pixels per composite tile (batched).
constexpr int kVariantK[kNumVariants] = {6, 6, 6, 6, 8, 8, 8, 8, 16, 16, 16, 16, 32, 32, 32, 32};
constexpr int kVariantCH[kNumVariants] = {64, 128, 256, 512, 64, 128, 256, 512,
                                          64, 128, 256, 512, 64, 128, 256, 512};

// One slot of the fence ring (kFenceSlots deep).  A page fence carries the
// error word of its page: submit_page copies the live word into err_flag[1 +
// slot] ahead of the event, so a later page's begin_page cannot erase it.
constexpr int kFenceSlots = 16;
struct Fence {
    hipEvent_t ev = nullptr;
    uint64_t id = 0;
    bool page = false;  // a page fence (else a transfer fence: no error word)
};

// Spans of one (path, 16-row band) item written by the batched span pass, in
// two arrays over the item index: a 16-byte head and a 128-byte row of counts.
// Lane l's disjoint sorted intervals [s, e) (AA x units) are the cnt[l] int2
// records starting at off + sum(cnt[0..l)).  off = -1: not stored; off = -2:
// skipped as occluded (the counts are not written).  fx0..fx1: pixel columns,
// within the path's drawn columns, that every one of the band's 64 AA rows
// covers fully; empty (fx0 > fx1) unless the path is opaque, so that the path
// paints them with exactly its colour.  cx0..cx1: the pixel columns the
// intervals touch within the drawn columns (empty when none), the item's
// extent for the composite tile bins, which read the heads alone.
struct SpanHead {
    long long off;
    int16_t fx0, fx1, cx0, cx1;
};
struct SpanCnt {
    uint16_t cnt[64];
};
// Sizing unit of the span workspace: a workspace of cap units holds cap
// heads followed by cap count rows.
struct SpanSlot {
    SpanHead h;
    SpanCnt c;
};
static_assert(sizeof(SpanHead) == 16 && sizeof(SpanCnt) == 128 && sizeof(SpanSlot) == 144, "span head layout");

// Everything a span item needs about its path and page, one 64-byte record
// per path (written by k_dispatch): a group of items is then described by two
// loads, the items and their path records.
struct SpanPath {
    long long e0, e1;          // edge records [e0, e1)
    int32_t y_min, y_max;      // scanner rows (AA space)
    int32_t row0, row1;        // drawn output rows
    int32_t px0, px1;          // drawn columns at dispatch (an unbalanced row widens the copy in sel)
    int32_t width, height;     // canvas
    int32_t px_hi;             // last column of the page clip
    int32_t occ_base, ctiles;  // the page's first composite tile, tiles per band
    uint32_t flags;            // bit 0 even-odd, bit 1 opaque AA paint, bits 8-15 a_input, bits 16-31 first clip column
};
static_assert(sizeof(SpanPath) == 64, "SpanPath layout");

// Composite work of one page: tiles per band and the index of its first
// (band, tile) item; pages are listed in the order their items are served.
struct CPage {
    int32_t page, tiles, base, pad;
};

}  // namespace rrdev

struct rr_hip_ctx {
    int device = 0;
    int n_cu = 0;
    int setup_bpc = 1;  // resident k_edge_setup blocks per CU
    hipStream_t stream = nullptr;
    std::string err;
    // workspace shared by both launch shapes
    rrdev::EdgeRec *recs = nullptr;
    size_t recs_cap = 0;
    rrdev::PathHdr *hdrs = nullptr;
    size_t hdrs_cap = 0;
    int2 *rows = nullptr;
    size_t rows_cap = 0;
    int4 *sel = nullptr;  // per path {first drawn row, last drawn row, first px, last px}
    size_t sel_cap = 0;
    int32_t *nb = nullptr;  // per path: 16-row bands it draws
    size_t nb_cap = 0;
    int2 *csum = nullptr;
    size_t csum_cap = 0;
    rrdev::DevPage *pages = nullptr;
    size_t pages_cap = 0;
    int32_t *slots = nullptr;
    size_t slots_cap = 0;  // in int32
    int32_t *spill = nullptr;
    long long spill_cap = 0;
    unsigned long long *spill_top = nullptr;
    int32_t *err_flag = nullptr;  // [0] live word of the open page, [1 + k] page fence slot k's copy
    // batched two-phase workspace
    long long *item_base = nullptr;  // per path: first span item (n_paths + 1)
    size_t item_base_cap = 0;
    long long *scan_tmp = nullptr;
    size_t scan_tmp_cap = 0;
    // device path ingestion (rr_xpath.hip): per point transformed position,
    // role, closing-line target, edge count and edge offset; an error flag
    double2 *xp_t = nullptr;
    size_t xp_t_cap = 0;
    uint8_t *xp_role = nullptr;
    size_t xp_role_cap = 0;
    long long *xp_close = nullptr;
    size_t xp_close_cap = 0;
    int32_t *xp_cnt = nullptr;
    size_t xp_cnt_cap = 0;
    long long *xp_eoff = nullptr;
    size_t xp_eoff_cap = 0;
    long long *xp_sidx = nullptr;  // segment index per point
    size_t xp_sidx_cap = 0;
    long long *xp_seg = nullptr;   // point per segment
    size_t xp_seg_cap = 0;
    double2 *xp_stage = nullptr;   // staged leaf points
    size_t xp_stage_cap = 0;
    uint8_t *xp_bkt = nullptr;     // work class per segment
    size_t xp_bkt_cap = 0;
    long long *xp_perm = nullptr;  // segments ordered by work class
    size_t xp_perm_cap = 0;
    int32_t *xp_ccnt = nullptr;    // segments per (class, block)
    size_t xp_ccnt_cap = 0;
    long long *xp_coff = nullptr;  // their exclusive scan
    size_t xp_coff_cap = 0;
    int32_t *xp_err = nullptr;
    size_t xp_err_cap = 0;
    hipEvent_t xp_ev[6] = {};  // walk, count, emit start/stop (profiling on)
    // image placement: grey + alpha per image pixel of an image converted ahead
    uint16_t *img_grey = nullptr;
    size_t img_grey_cap = 0;
    double xp_ms[3] = {0.0, 0.0, 0.0};
    int4 *sitems = nullptr;  // span-pass work items (path, first row, bands, first span item)
    size_t sitems_cap = 0;
    rrdev::SpanSlot *shdr = nullptr;  // heads, then count rows (see SpanSlot)
    size_t shdr_cap = 0;
    int2 *spans = nullptr;
    size_t spans_cap = 0;
    int32_t *band_cnt = nullptr;  // per batch-global band: +1/-1 marks of the paths' band ranges (n_bands + 2)
    size_t band_cnt_cap = 0;
    long long *band_run = nullptr;  // exclusive prefix of the marks: band_run[i + 1] = paths drawing band i
    size_t band_run_cap = 0;
    long long *band_base = nullptr;  // exclusive prefix of the per-band path counts (n_bands + 1)
    size_t band_base_cap = 0;
    int2 *band_list = nullptr;  // per band, in path order: (path, span item)
    size_t band_list_cap = 0;
    rrdev::CPage *cpages = nullptr;
    size_t cpages_cap = 0;
    int32_t *occ = nullptr;  // per composite tile item: highest opaque path covering it fully, or -1
    size_t occ_cap = 0;
    int32_t *pstate = nullptr;  // per path: bit 0 set once an unbalanced row widened the path
    size_t pstate_cap = 0;
    rrdev::SpanPath *spath = nullptr;  // per path: the span pass's record of the path
    size_t spath_cap = 0;
    int4 *ovq = nullptr;  // bands the main span pass deferred to the wide pass
    size_t ovq_cap = 0;
    int32_t *bin_cnt = nullptr;  // per composite tile: entries of its bin
    size_t bin_cnt_cap = 0;
    long long *bin_off = nullptr;  // per composite tile: first entry of its bin (n + 1)
    size_t bin_off_cap = 0;
    int4 *bins = nullptr;  // composite tile bins: {span item, path, full-cover columns, attribute bytes}
    size_t bins_cap = 0;
    long long batch_paths = 0;  // paths of the last batched recording
    long long batch_bins = 0;  // bin entries of the last batched recording (the sizing pass's while capturing)
    // dispatcher
    float *feat = nullptr;  // per path: kFeatures floats
    size_t feat_cap = 0;
    uint8_t *route_cls = nullptr;  // per path: class chosen (RR_ROUTE_*)
    size_t route_cls_cap = 0;
    int32_t *n_work = nullptr;  // per path: span-pass work items
    size_t n_work_cap = 0;
    long long *work_base = nullptr;  // exclusive prefix of n_work (n_paths + 1)
    size_t work_base_cap = 0;
    float *mlp = nullptr;  // dispatcher MLP parameters (kMlpParams floats) or unused
    bool mlp_loaded = false;
    // [0] span items served, [1] composite items served, [2] span pool top,
    // [3] intervals stored, [4] (path, tile) composites, [5] tiles touched
    unsigned long long *counters = nullptr;
    unsigned long long *queues = nullptr;  // kQueueWords
    // profiling event pairs of the last batched recording: phase A [a0, a1), phase B b
    size_t prof_a0 = 0, prof_a1 = 0, prof_b = 0;
    int64_t batch_items = 0, batch_citems = 0, batch_reruns = 0, batch_work = 0;
    bool batch_valid = false;
    // graph capture of a batched recording: while capturing, nothing syncs or
    // uploads (the tables and sizes come from the eager sizing pass before it)
    bool capturing = false, size_only = false;
    uint64_t ws_gen = 0;  // bumped whenever a workspace array is reallocated
    int span_waves[rrdev::kNumVariants] = {};
    int compose_waves[rrdev::kNumVariants] = {};
    int compose_waves3[rrdev::kNumVariants] = {};  // the Rgb8 (three-plane) composite pass
    int wide_waves = 1;  // persistent grid of the wide span pass
    std::vector<rrdev::CPage> host_cpages;
    // page state
    hipEvent_t t0 = nullptr, t1 = nullptr;
    bool page_open = false, page_timed = false;
    // Plain launches go to kPlainStreams side streams (created on first use):
    // a path whose drawn rows meet those of an earlier path of its page runs
    // after it (same stream, or an event wait); other paths run concurrently.
    // Each stream wants a hardware queue of its own (GPU_MAX_HW_QUEUES of at
    // least kPlainStreams + 2: these, the context stream and HIP's own).
    static constexpr int kPlainStreams = 20;
    int plain_streams = kPlainStreams;  // side streams in use (from GPU_MAX_HW_QUEUES)
    rrdev::FillArgs *plain_args = nullptr;  // device copy of the plain launches' shared arguments
    // batched recordings: the band lists are built on `aux` beside the span
    // pass (fork after the band offsets, join before the tile bins)
    hipStream_t aux = nullptr;
    hipEvent_t aux_fork = nullptr, aux_join = nullptr;
    static constexpr int kPlainCellRows = 4;  // row granularity of the overlap test
    hipStream_t side[kPlainStreams] = {};
    hipEvent_t side_ev[kPlainStreams] = {};  // dependency and join points of each side stream
    hipEvent_t side_start = nullptr;         // the context stream's work before the launches
    std::vector<int8_t> cell_stream;         // per row cell of the page: last side stream drawing it
    double last_ms = 0.0;
    int64_t last_launches = 0;
    rrdev::Fence fences[rrdev::kFenceSlots];
    // sizes of the buffers rr_hip_alloc handed out (device_buffer_len)
    std::unordered_map<const void *, size_t> buf_len;
    uint64_t next_fence = 1;
    // this device's address of the static scan scratch (rr_hip_record_scan_static)
    uint32_t *static_scratch = nullptr;
    // per-launch profiling
    bool profiling = false;
    std::vector<hipEvent_t> prof_ev;  // pairs
    size_t prof_used = 0;
    double last_fill_ms = 0.0;
    int64_t fill_launches = 0;
    std::vector<rrdev::DevPage> host_pages;
    // optional fill timing trace (caller-owned device buffer)
    longlong4 *trace = nullptr;
    long long trace_cap = 0;
    unsigned long long *trace_top = nullptr;
};

namespace rrint {
hipStream_t stream(rr_hip_ctx *c);
int device(rr_hip_ctx *c);
bool page_open(rr_hip_ctx *c);
int32_t fail(rr_hip_ctx *c, int32_t code, const std::string &msg);
// Start/stop event pair for the next kernel when per-launch profiling is on,
// nullptr otherwise; *ok is false when the pool could not grow.
hipEvent_t *prof_pair(rr_hip_ctx *c, bool *ok);
void count_launch(rr_hip_ctx *c);

// Grows a device workspace array to hold at least `need` elements (doubling,
// contents not kept; the stream is drained before the old array is freed).
template <typename T>
int32_t ensure(rr_hip_ctx *c, T **buf, size_t *cap, size_t need) {
    if (need <= *cap) return RR_OK;
    if (c->capturing) return fail(c, RR_E_INVARIANT, "workspace would grow during graph capture");
    ++c->ws_gen;
    size_t n = *cap ? *cap : 1024;
    while (n < need) n *= 2;
    hipError_t e = hipSuccess;
    if (*buf) {
        e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipFree(*buf);
        if (e != hipSuccess) return fail(c, RR_E_DEVICE, std::string("workspace free: ") + hipGetErrorString(e));
    }
    *buf = nullptr;
    *cap = 0;
    e = hipMalloc(reinterpret_cast<void **>(buf), n * sizeof(T));
    if (e != hipSuccess) return fail(c, RR_E_DEVICE, std::string("workspace alloc: ") + hipGetErrorString(e));
    *cap = n;
    return RR_OK;
}

// Records the batched two-phase fill + composite of b (rr_batch.hip).  A holds
// the per-batch arguments already set up by the caller; the edge-setup and
// chunk-summary kernels have been recorded.
int32_t record_batched(rr_hip_ctx *c, const rr_fill_pages &b, const rrdev::FillArgs &A, int k_entries,
                       int chunk_px, int variant);
// Occupancy-derived persistent wave counts of the batched kernels.
int32_t init_batched(rr_hip_ctx *c);
// Exclusive scan of n int32 counts into out[0..n), out[n] = total (rr_batch.hip).
int32_t scan_counts(rr_hip_ctx *c, const int32_t *in, long long *out, long long n);
// Zero n 64-bit words with a kernel (safe inside graph capture).
int32_t zero_u64(rr_hip_ctx *c, unsigned long long *p, long long n);
int32_t fill_u64(rr_hip_ctx *c, unsigned long long *p, long long n, unsigned long long value);
int32_t fill_i32(rr_hip_ctx *c, int32_t *p, long long n, int32_t value);
// Clears the bits of `mask` in the context's live error word, keeping the
// others (a kernel, so that it is ordered on the stream and capturable).
int32_t clear_err_bits(rr_hip_ctx *c, int32_t mask);
int32_t set_err_bits(rr_hip_ctx *c, int32_t bits);
// Looks up (once per device, that device current) the address of the static
// scan scratch and keeps it in the context (rr_batch.hip).
int32_t init_static_scratch(rr_hip_ctx *c);
// Records rgb_to_gray of an Rgb8 bitmap into a Gray8 one (rr_image.hip).
int32_t rgb_to_gray(rr_hip_ctx *c, const uint8_t *rgb, int64_t rgb_stride, uint8_t *gray, int64_t gray_stride,
                    int32_t width, int32_t height);
// Records the edge setup + fill of b (rr_fill.hip); used by graph capture.
int32_t record_pages_internal(rr_hip_ctx *c, const rr_fill_pages &b);
}  // namespace rrint
