// rr_batch.hip -- the batched (many pages, one recording) fill + composite,
// in two phases, and the device exclusive scan it is built on.
//
// Phase A (spans): one work item per (path, 16-row band) the path draws.
//   Persistent waves pull items in any order; each builds the band's 64 AA
//   row entry lists and winding walk exactly as the plain route does
//   (scanner/mod.rs:511-642, render_aa_line :393-467) and stores the result,
//   per AA row, as disjoint sorted intervals [s, e) in a span pool.  Items are
//   independent, so the load balances over the whole batch.
// Phase B (composite): one work item per (page, band, x tile of CH pixels).
//   The wave keeps the tile's 16 x CH canvas bytes in LDS, walks the band's
//   path list in path order, turns each path's stored intervals into the
//   4x4 coverage bitmap of the tile and composites source-over in LDS
//   (fill/mod.rs:344-401, pipe/aa.rs:192-208, simd/composite.rs:56-114), then
//   writes the tile back once.  Tiles are disjoint, so per-pixel path order
//   holds without inter-wave ordering.
// The per-band path lists come from a device stable bucketing of the span
// items: count per band, exclusive scan, and an in-order ballot sweep per band.
//
// The scan (three kernels: tile sums, scan of the sums, tile rescan) replaces
// the reference's Blelloch scan (crates/gpu/kernels/blelloch_scan.cu) with
// wave64 shuffle scans: 256 threads x 4 elements = 1024-element tiles.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "rr_hip.h"
#include "rr_internal.h"
#include "rr_mlp.h"

using namespace rrdev;

namespace {

// ---------------------------------------------------------------------------
// exclusive scan
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256, kScanPer = 4, kScanTile = kScanThreads * kScanPer;

template <typename T>
__device__ __forceinline__ T wave_incl(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread (256 threads); *total gets
// the block sum.  ws: 4 LDS slots.
template <typename T>
__device__ __forceinline__ T block_excl(T v, T *ws, T *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const T inc = wave_incl(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / 64; ++k) {
        if (k < w) base += ws[k];
        tot += ws[k];
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const Tin *__restrict__ in, long long n,
                                                              Tout *__restrict__ sums) {
    __shared__ Tout ws[kScanThreads / 64];
    const long long i0 = static_cast<long long>(blockIdx.x) * kScanTile + kScanPer * threadIdx.x;
    Tout v = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
        if (i0 + k < n) v += static_cast<Tout>(in[i0 + k]);
    Tout tot;
    (void)block_excl(v, ws, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One block: exclusive scan of the tile sums in place, 1024 at a time with a
// carry; *total (optional) receives the grand total.
template <typename Tout>
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(Tout *__restrict__ sums, long long nb, Tout *total) {
    __shared__ Tout ws[kScanThreads / 64];
    Tout carry = 0;
    for (long long base = 0; base < nb; base += kScanTile) {
        const long long i0 = base + kScanPer * threadIdx.x;
        Tout v[kScanPer];
        Tout s = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = i0 + k < nb ? sums[i0 + k] : static_cast<Tout>(0);
            s += v[k];
        }
        Tout tot;
        Tout run = carry + block_excl(s, ws, &tot);
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (i0 + k < nb) sums[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (total && threadIdx.x == 0) *total = carry;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const Tin *in, Tout *out, long long n,
                                                            const Tout *__restrict__ sums) {
    __shared__ Tout ws[kScanThreads / 64];
    const long long i0 = static_cast<long long>(blockIdx.x) * kScanTile + kScanPer * threadIdx.x;
    Tout v[kScanPer];
    Tout s = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = i0 + k < n ? static_cast<Tout>(in[i0 + k]) : static_cast<Tout>(0);
        s += v[k];
    }
    Tout tot;
    Tout run = sums[blockIdx.x] + block_excl(s, ws, &tot);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (i0 + k < n) out[i0 + k] = run;
        run += v[k];
    }
}

__global__ void k_and_i32(int32_t *p, int32_t keep) { atomicAnd(p, keep); }
__global__ void k_or_i32(int32_t *p, int32_t bits) { atomicOr(p, bits); }

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ p, long long n, T value) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += static_cast<long long>(gridDim.x) * 256)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// span items and per-band path lists
// ---------------------------------------------------------------------------
struct BatchArgs {
    FillArgs F;
    const SpanPath *spath;
    int4 *sel;
    const long long *item_base;
    long long n_paths;
    const int4 *sitems;  // work items: (path, first row, bands, first span item)
    long long sitems_cap;
    const long long *work_base;  // n_paths + 1: work items per path, exclusive prefix
    long long shdr_cap;
    SpanHead *shead;  // shdr_cap heads, then shdr_cap count rows (span_cnt)
    int2 *spans;
    long long spans_cap;
    const long long *band_base;
    const int2 *band_list;
    const CPage *cpages;
    int32_t n_cpages;
    unsigned long long *counters;
    int32_t *occ;  // per composite tile: highest opaque path index covering it fully
    int32_t chunk_px, chunk_shift;  // composite tile width, a power of two, and its log2
    int32_t *pstate;  // per path: bit 0 = widened by an unbalanced row, bit 1 = a band was built
    int4 *ovq;        // deferred bands: {path, first row, span item, 0}
    long long ovq_cap;
    const long long *bin_off;  // per composite item: first entry of its tile bin (n + 1)
    const int4 *bins;          // tile bin entries: {span item, path, full-cover columns, attribute bytes}
    unsigned long long *queues;  // sliced queue counters: span pass, then composite pass
};

// A field of the kernel's BatchArgs argument, read from the kernel-argument
// segment where it is used (one scalar load) instead of being held in SGPRs
// for the whole kernel: for what the span pass needs rarely (error word,
// counters, caps, the deferred queue).  Only for kernels whose first argument
// is the BatchArgs (k_spans, k_spans_wide).
#define RR_KARG_AS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ T karg_at(size_t off) {
    const RR_KARG_AS char *base = (const RR_KARG_AS char *)(__builtin_amdgcn_kernarg_segment_ptr());
    return *reinterpret_cast<const volatile RR_KARG_AS T *>(base + off);
}
#define RR_COLD(field) karg_at<decltype(BatchArgs{}.field)>(offsetof(BatchArgs, field))

// Span pass modes: every item, counting drawn widths and skipping occluded
// items (skipping only when no per-path coverage output is asked for); the
// same without counting (a rerun after growing the interval pool); every item
// without skipping (the sizing pass before a capture); or (wide pass only)
// the skipped items of paths an unbalanced row widened after they were skipped.
enum { SPAN_ALL = 0, SPAN_ALL_NODRAW = 1, SPAN_ALL_NOSKIP = 2, SPAN_FIXUP = 3 };
// Count row of span item it (after the shdr_cap heads).
__device__ __forceinline__ SpanCnt *span_cnt(const BatchArgs &B, long long it) {
    return reinterpret_cast<SpanCnt *>(B.shead + B.shdr_cap) + it;
}
constexpr long long kSkipped = -2;  // SpanHead::off of an item skipped as occluded

// The dispatcher, in-kernel ahead of the fill: the class of every path (the
// caller's route, or for RR_ROUTE_AUTO the MLP's smallest predicted span-pass
// time among the valid classes, or the static rule without an MLP) and the
// span-pass work items it takes.  One thread per path.
__global__ void __launch_bounds__(256) k_dispatch(const rr_path_attr *__restrict__ attrs, const int32_t *__restrict__ nb,
                                                  const int4 *__restrict__ sel, const float *__restrict__ feat,
                                                  const float *__restrict__ mlp, long long n_paths,
                                                  uint8_t *__restrict__ route_cls, int32_t *__restrict__ n_work,
                                                  const PathHdr *__restrict__ hdrs, const DevPage *__restrict__ pages,
                                                  const int64_t *__restrict__ edge_off, SpanPath *__restrict__ spath) {
    __shared__ float P[kMlpParams];
    if (mlp) {
        for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) P[i] = mlp[i];
        __syncthreads();
    }
    const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n_paths) return;
    const int32_t bands = nb[p];
    const rr_path_attr at = attrs[p];
    const int4 s = sel[p];
    {
        const PathHdr h = hdrs[p];
        const DevPage &pg = pages[h.page];
        SpanPath sp;
        sp.e0 = edge_off[p];
        sp.e1 = edge_off[p + 1];
        sp.y_min = h.y_min;
        sp.y_max = h.y_max;
        sp.row0 = s.x;
        sp.row1 = s.y;
        sp.px0 = s.z;
        sp.px1 = s.w;
        sp.width = pg.width;
        sp.height = pg.height;
        sp.px_hi = pg.px_hi;
        sp.occ_base = pg.cbase;
        sp.ctiles = pg.ctiles;
        // bit 1: an occluder candidate -- an opaque paint on the AA path, which
        // leaves exactly its colour under full coverage
        const bool opaque = at.a_input == 255 && !(at.eo & kFillGeneral);
        sp.flags = (at.eo & RR_FILL_EVEN_ODD) | (opaque ? 2u : 0u) | (static_cast<uint32_t>(at.a_input) << 8) |
                   (static_cast<uint32_t>(min(max(pg.px_lo, 0), 0xffff)) << 16);
        spath[p] = sp;
    }
    int cls = at.route;
    auto valid = [&](int c) {
        return c == RR_ROUTE_SPAN || c == RR_ROUTE_CPU || (c == RR_ROUTE_TILE16 && bands <= 1) ||
               (c == RR_ROUTE_TILE32 && bands <= 2) || (c == RR_ROUTE_TILE64 && bands <= 4);
    };
    if (cls == RR_ROUTE_AUTO) {
        if (mlp) {
            float out[kClasses];
            mlp_forward(P, feat + p * kFeatures, out);
            cls = RR_ROUTE_SPAN;
            for (int c = 1; c < kClasses; ++c)
                if (valid(c) && out[c] < out[cls]) cls = c;
        } else {
            // static rule: the reference's bbox area (gpu_ops.rs:52-84) below
            // 256 px, or no finite bbox -> CPU (GPU_TILE/AA_FILL_THRESHOLD,
            // crates/gpu/src/lib.rs:112,120; gpu_ops.rs:242,289), else the
            // smallest tile that holds the path's bands, else span
            const int32_t area = hdrs[p].ref_area;
            if (area < 256)
                cls = RR_ROUTE_CPU;
            else if (bands <= 1)
                cls = RR_ROUTE_TILE16;
            else if (bands <= 2)
                cls = RR_ROUTE_TILE32;
            else if (bands <= 4)
                cls = RR_ROUTE_TILE64;
            else
                cls = RR_ROUTE_SPAN;
        }
    }
    if (cls > RR_ROUTE_CPU || !valid(cls)) cls = RR_ROUTE_SPAN;
    route_cls[p] = static_cast<uint8_t>(cls);
    const bool tile = cls == RR_ROUTE_TILE16 || cls == RR_ROUTE_TILE32 || cls == RR_ROUTE_TILE64;
    n_work[p] = bands == 0 ? 0 : (tile ? 1 : bands);
}

// Work items of path p (one per band for SPAN, one for all bands for a TILE
// class) and the per-band path counts as marks: +1 at the path's first
// batch-global band, -1 one past its last (the next page's first band or the
// end), so that the prefix sum of the marks counts the paths drawing each
// band.  One wave per path.
__global__ void __launch_bounds__(256) k_expand(const int4 *__restrict__ sel, const int32_t *__restrict__ nb,
                                                const int32_t *__restrict__ n_work, const PathHdr *__restrict__ hdrs,
                                                const DevPage *__restrict__ pages,
                                                const long long *__restrict__ item_base,
                                                const long long *__restrict__ work_base, long long n_paths,
                                                int4 *__restrict__ sitems, long long sitems_cap,
                                                int32_t *__restrict__ band_marks, int32_t *__restrict__ err) {
    const long long p = static_cast<long long>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= n_paths) return;
    const int32_t n = nb[p];
    if (n == 0) return;
    const long long base = item_base[p], wbase = work_base[p];
    const int32_t b0 = sel[p].x / kBandRows;
    const long long gb0 = static_cast<long long>(pages[hdrs[p].page].band_base) + b0;
    const int32_t nw = n_work[p];
    for (int32_t k = lane; k < n; k += 64) {
        if (k < nw) {
            const long long it = wbase + k;
            if (it >= sitems_cap) {
                atomicOr(err, 2);
            } else if (nw == n) {
                sitems[it] = make_int4(static_cast<int32_t>(p), kBandRows * (b0 + k), 1, static_cast<int32_t>(base + k));
            } else {
                sitems[it] = make_int4(static_cast<int32_t>(p), kBandRows * b0, n, static_cast<int32_t>(base));
            }
        }
    }
    if (lane == 0) {
        atomicAdd(band_marks + gb0, 1);
        atomicAdd(band_marks + gb0 + n, -1);
    }
}

// Band gb's list of (path, span item) in path order: an in-order sweep of its
// page's paths, compacted with ballots.  One wave per band.
__global__ void __launch_bounds__(64) k_band_lists(const int4 *__restrict__ sel, const DevPage *__restrict__ pages,
                                                   int32_t n_pages, const long long *__restrict__ item_base,
                                                   const long long *__restrict__ band_base,
                                                   int2 *__restrict__ band_list, long long list_cap,
                                                   int32_t *__restrict__ err) {
    const long long gb = blockIdx.x;
    const int lane = threadIdx.x;
    int32_t lo = 0, hi = n_pages - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (pages[mid].band_base <= gb)
            lo = mid;
        else
            hi = mid - 1;
    }
    const DevPage &pg = pages[lo];
    const int32_t band = static_cast<int32_t>(gb - pg.band_base);
    const int32_t Y0 = kBandRows * band, Y1 = Y0 + kBandRows - 1;
    long long w = band_base[gb];
    const long long wend = band_base[gb + 1];
    const long long p_end = pg.path_end;
    for (long long base = pg.path_begin; base < p_end; base += 64) {
        const long long q = base + lane;
        bool ov = false;
        int4 s = make_int4(1, 0, 1, 0);
        if (q < p_end) {
            s = sel[q];
            ov = s.x <= s.y && s.x <= Y1 && s.y >= Y0;
        }
        const unsigned long long m = __ballot(ov);
        if (ov) {
            const long long pos = w + __popcll(m & ((1ull << lane) - 1ull));
            if (pos < wend && pos < list_cap)
                band_list[pos] = make_int2(static_cast<int32_t>(q),
                                           static_cast<int32_t>(item_base[q] + (band - s.x / kBandRows)));
            else
                atomicOr(err, 4);
        }
        w += __popcll(m);
    }
}

// Composite tile bins: for every (band, tile) composite item, the span items
// of the paths whose stored intervals touch the tile, in path order.  A
// band's path list (already in path order) is cut into kBinSegs consecutive
// segments, one wave each, so long lists are swept in parallel; each visible
// item's head gives its column extent.  The counting pass writes the count of
// every (bin, segment), indexed (cbase + band * ctiles + tile) * kBinSegs +
// segment, so that one exclusive scan yields each segment's first position
// inside its bin, and bin ci spans [bin_off[ci * kBinSegs],
// bin_off[(ci + 1) * kBinSegs]); the placing pass puts each item at its
// segment's offset + its rank among the segment's earlier items in that bin.
// Items skipped as occluded, items without intervals and (item, tile) pairs
// that a later opaque path covers fully (occ[tile] above the path: the
// composite starts after that path) are left out.
constexpr int kMaxTiles = 512;  // a 32767-pixel page at 64-pixel tiles
constexpr int kBinSegs = 4;
template <bool PLACE>
__global__ void __launch_bounds__(64) k_tile_bins(const DevPage *__restrict__ pages, int32_t n_pages,
                                                  const long long *__restrict__ band_base,
                                                  const int2 *__restrict__ band_list, const SpanHead *__restrict__ shead,
                                                  long long shdr_cap, int32_t chunk_px, const int32_t *__restrict__ occ,
                                                  const rr_path_attr *__restrict__ attrs,
                                                  int32_t *__restrict__ bin_cnt, const long long *__restrict__ bin_off,
                                                  int4 *__restrict__ bins, long long bins_cap,
                                                  int32_t *__restrict__ err) {
    __shared__ long long run[kMaxTiles];
    __shared__ int32_t cover[kMaxTiles];
    const long long gb = blockIdx.x / kBinSegs;
    const int32_t seg = static_cast<int32_t>(blockIdx.x % kBinSegs);
    const int lane = threadIdx.x;
    int32_t lo = 0, hi = n_pages - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (pages[mid].band_base <= gb)
            lo = mid;
        else
            hi = mid - 1;
    }
    const DevPage &pg = pages[lo];
    const int32_t band = static_cast<int32_t>(gb - pg.band_base);
    const int32_t T = min(pg.ctiles, kMaxTiles);
    const long long ci0 = static_cast<long long>(pg.cbase) + static_cast<long long>(band) * pg.ctiles;
    for (int32_t t = lane; t < T; t += 64) {
        run[t] = PLACE ? bin_off[(ci0 + t) * kBinSegs + seg] : 0ll;
        cover[t] = occ[ci0 + t];
    }
    __syncthreads();
    // this wave's segment of the band's list, in whole 64-entry chunks
    const long long b0 = band_base[gb], b1 = band_base[gb + 1];
    const long long seg_len = ((b1 - b0 + kBinSegs - 1) / kBinSegs + 63) & ~63ll;
    const long long l0 = min(b1, b0 + seg * seg_len), l1 = min(b1, l0 + seg_len);
    for (long long lb = l0; lb < l1; lb += 64) {
        const long long q = lb + lane;
        int2 ent = make_int2(0, 0);
        int32_t ta = INT32_MAX, tb = -1, fx = 0;
        if (q < l1) {
            ent = band_list[q];
            if (ent.y >= 0 && ent.y < shdr_cap) {
                const int4 h = *reinterpret_cast<const int4 *>(shead + ent.y);
                fx = h.z;
                const long long off = static_cast<long long>((static_cast<uint64_t>(static_cast<uint32_t>(h.y)) << 32) |
                                                             static_cast<uint32_t>(h.x));
                const int32_t cx0 = static_cast<int16_t>(h.w & 0xffff), cx1 = static_cast<int16_t>(h.w >> 16);
                if (off != kSkipped && cx0 <= cx1) {
                    ta = cx0 / chunk_px;
                    tb = min(cx1 / chunk_px, T - 1);
                }
            } else {
                atomicOr(err, 16);
            }
        }
        const int32_t u0 = wave_min(ta), u1 = wave_max(tb);
        for (int32_t t = u0; t <= u1; ++t) {
            const bool in = ta <= t && t <= tb && cover[t] <= ent.x;
            const uint64_t m = __ballot(in);
            if (m == 0) continue;
            if (PLACE && in) {
                const long long pos = run[t] + __popcll(m & ((1ull << lane) - 1ull));
                if (pos < bins_cap)
                    bins[pos] = make_int4(ent.y, ent.x, fx, static_cast<int32_t>(reinterpret_cast<const uint32_t *>(attrs)[ent.x]));
                else
                    atomicOr(err, 64);
            }
            __syncthreads();
            if (lane == 0) run[t] += __popcll(m);
            __syncthreads();
        }
    }
    if (!PLACE)
        for (int32_t t = lane; t < T; t += 64) bin_cnt[(ci0 + t) * kBinSegs + seg] = static_cast<int32_t>(run[t]);
}

// ---------------------------------------------------------------------------
// phase A: spans of one (path, band)
// ---------------------------------------------------------------------------
template <int K>
struct SpanLds {
    int4 recA[1][64], recB[1][64];  // the chunk's EdgeRecs as their two 16-byte halves
    unsigned long long smask[64];
    int32_t ex0[K * 64];
    int32_t ex1[K * 64];
    int32_t ec[K * 64];
};

// Per-wave sub-allocator over the span pool: one global atomic per chunk of
// kSpanChunk records instead of one per item.
constexpr long long kSpanChunk = 8192;
struct SpanCursor {
    unsigned long long next = 0, end = 0;
    uint32_t stored = 0;   // intervals stored by this wave
    uint32_t skipped = 0;  // bands skipped as occluded
    int32_t marked = -1;             // last path marked as having a visible band
};

// A span item's path as wave-uniform values (unpacked from its SpanPath).
struct ItemPath {
    int64_t p, e0, e1;
    int32_t y_min, y_max, row0, row1, px0, px1, width, height, px_lo, px_hi, occ_base, ctiles;
    int32_t eo, opaque;
};

// First 16 bytes of a span header (offset, full-cover and touched columns),
// one store by lane 0.
__device__ __forceinline__ void hdr_head(SpanHead *H, long long off, int32_t fx0, int32_t fx1, int32_t cx0,
                                         int32_t cx1) {
    if (threadIdx.x == 0)
        *reinterpret_cast<int4 *>(H) =
            make_int4(static_cast<int32_t>(off & 0xffffffffll), static_cast<int32_t>(off >> 32),
                      static_cast<int32_t>((static_cast<uint32_t>(fx0) & 0xffffu) | (static_cast<uint32_t>(fx1) << 16)),
                      static_cast<int32_t>((static_cast<uint32_t>(cx0) & 0xffffu) | (static_cast<uint32_t>(cx1) << 16)));
}

// The intervals of one (path, band) from its entry lists X0/X1/C (LDS or the
// wave's global slot): sort and walk, drawn widths, fully covered columns,
// store, header, occlusion publish.
// Sort and walk of a row holding at most two entries, in registers (what
// sort_and_walk does for n <= 2): its one interval [*s, *e), if any.
template <typename Ptr>
__device__ __forceinline__ int walk_short(int n, Ptr X0, Ptr X1, Ptr C, int32_t mrule, int32_t aa_width,
                                          bool *unbalanced, int32_t *s, int32_t *e) {
    const int lane = threadIdx.x;
    if (n == 0) return 0;
    const int32_t a0 = X0[lane], b0 = X1[lane], c0 = C[lane];
    int32_t start = a0, end = aa_width;
    if (n == 2) {
        const int32_t a1 = X0[64 + lane], b1 = X1[64 + lane], c1 = C[64 + lane];
        // stable order: the second entry goes first only when strictly left
        const bool swap = a1 < a0;
        start = swap ? a1 : a0;
        const int32_t x1 = swap ? b0 : b1;
        const int32_t count = static_cast<int32_t>(static_cast<uint32_t>(c0) + static_cast<uint32_t>(c1));
        if ((count & mrule) == 0)
            end = x1 == INT32_MAX ? INT32_MAX : x1 + 1;
        else
            *unbalanced = true;
    } else {
        *unbalanced = true;  // one state-changing entry never closes
    }
    const int32_t bx0 = start > 0 ? start : 0;
    int32_t bx1 = end < aa_width ? end : aa_width;
    if (bx1 < 0) bx1 = 0;
    if (bx0 >= bx1) return 0;
    *s = bx0;
    *e = bx1;
    return 1;
}

// One interval per lane held in a register, indexed like an interval list
// (any index is slot 0): span_finish over the short walk's result.
struct RegIval {
    int32_t v;
    __device__ __forceinline__ int32_t &operator[](long long) { return v; }
    __device__ __forceinline__ int32_t operator[](long long) const { return v; }
};

// nint, unb: the walk's interval count of this lane's row (in X0/X1 slots)
// and whether the row is unbalanced.
// MAXN: 1 when every lane holds at most one interval (RegIval lists), 0 for
// any count.
template <int MAXN, typename Ptr>
__device__ __forceinline__ void span_finish(const BatchArgs &B, const ItemPath &P, int32_t Y0, long long it,
                                            const int32_t *occ_row, Ptr &X0, Ptr &X1, int nint_in, bool unb,
                                            bool count_draw, bool occlusion, uint64_t vmask, SpanCursor &cur) {
    const int nint = MAXN == 1 ? min(nint_in, 1) : nint_in;
    const int lane = threadIdx.x;
    const int32_t y = Y0 + (lane >> 2), j = lane & 3;
    // an unbalanced row spans to the canvas edge: widen the path's columns;
    // its items skipped as occluded are redone by the fixup pass
    const bool any_unb = __ballot(unb) != 0;
    if (any_unb && lane == 0) {
        atomicMax(reinterpret_cast<int32_t *>(RR_COLD(sel) + P.p) + 3, P.px_hi);
        if ((atomicOr(RR_COLD(pstate) + P.p, 1) & 1) == 0) atomicAdd(RR_COLD(counters) + kCntWidened, 1ull);
    }
    const int32_t px1 = any_unb ? P.px_hi : P.px1;  // last drawn column of this item
    if (count_draw && RR_COLD(F.cov_count)) {
        // draw_aa_line range of each output row (scanner/mod.rs:459-466)
        int32_t smin = nint ? X0[lane] : INT32_MAX;
        int32_t emax = nint ? X1[(nint - 1) * 64 + lane] : -1;
        smin = min(smin, dpp_i32<kDppQuadXor1>(smin));
        smin = min(smin, dpp_i32<kDppQuadXor2>(smin));
        emax = max(emax, dpp_i32<kDppQuadXor1>(emax));
        emax = max(emax, dpp_i32<kDppQuadXor2>(emax));
        int32_t w_draw = 0;
        if (j == 0 && y >= P.row0 && y <= P.row1 && smin < emax) {
            const int32_t x0 = max(smin / 4, P.px_lo), x1 = min((emax - 1) / 4, P.px_hi);
            if (x0 <= x1) w_draw = x1 - x0 + 1;
        }
        const int32_t tot = __builtin_amdgcn_readlane(wave_incl_add(w_draw), 63);
        if (lane == 0 && tot)
            atomicAdd(reinterpret_cast<unsigned long long *>(RR_COLD(F.cov_count) + 2 * P.p + 1),
                      static_cast<unsigned long long>(tot));
    }
    // columns fully covered on all 64 AA rows: per row the longest interval's
    // whole pixels, intersected over the band (a subset of the exact set)
    int32_t fx0 = INT32_MAX, fx1 = INT32_MIN;
    if (nint > 0) {
        int32_t best = -1;
        for (int k = 0; k < nint; ++k) {
            const int32_t a = X0[k * 64 + lane], b = X1[k * 64 + lane];
            if (b - a > best) {
                best = b - a;
                fx0 = (a + 3) >> 2;
                fx1 = (b >> 2) - 1;
            }
        }
    }
    fx0 = max(wave_max(fx0), P.px0);
    fx1 = min(wave_min(fx1), px1);
    if (!P.opaque || fx0 > fx1) {
        fx0 = 1;
        fx1 = 0;
    }
    // Intervals kept: those that reach, within the drawn columns, a composite
    // tile not yet covered by a later opaque path (bit i of vmask: tile
    // px0 / chunk_px + i; tiles past the mask count as visible).  The others
    // would never be composited.  An unbalanced item keeps everything.  Kept
    // intervals are clipped to the drawn columns.
    const uint64_t vm = any_unb ? ~0ull : vmask;
    const int32_t sh = B.chunk_shift, t0 = P.px0 >> sh;
    int keep = 0;
    int32_t cx0 = INT32_MAX, cx1 = INT32_MIN;
    for (int k = 0; k < nint; ++k) {
        const int32_t a = X0[k * 64 + lane], b = X1[k * 64 + lane];
        const int32_t pa = max(a >> 2, P.px0), pb = min((b - 1) >> 2, px1);
        if (pa > pb) continue;
        const int32_t ta = (pa >> sh) - t0, tb = (pb >> sh) - t0;
        const uint64_t bits = tb >= 63 ? ~0ull << ta : ((2ull << tb) - 1ull) & (~0ull << ta);
        if (tb < 64 && (vm & bits) == 0) continue;
        // compact in place, clipped to the drawn columns (the composite pass
        // then clips to its tile only)
        X0[keep * 64 + lane] = max(a, 4 * P.px0);
        X1[keep * 64 + lane] = min(b, 4 * (px1 + 1));
        ++keep;
        cx0 = min(cx0, pa);
        cx1 = max(cx1, pb);
    }
    cx0 = wave_min(cx0);
    cx1 = wave_max(cx1);
    if (cx0 > cx1) {
        cx0 = 1;
        cx1 = 0;
    }
    // store the kept intervals: lane-major, one int2 per interval
    const int32_t incl = wave_incl_add(keep);
    const int32_t excl = incl - keep;
    const int32_t total = __builtin_amdgcn_readlane(incl, 63);
    if (cur.next + static_cast<unsigned long long>(total) > cur.end) {
        const unsigned long long want =
            static_cast<unsigned long long>(total > kSpanChunk ? total : kSpanChunk);
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(RR_COLD(counters) + kCntPoolTop, want);
        base = first_lane_u64(base);
        cur.next = base;
        cur.end = base + want;
    }
    const unsigned long long off = cur.next;
    cur.next += static_cast<unsigned long long>(total);
    const bool fits = static_cast<long long>(off) + total <= RR_COLD(spans_cap);
    cur.stored += fits ? static_cast<uint32_t>(total) : 0u;
    if (fits) {
        int2 *dst = B.spans + off + excl;
        for (int k = 0; k < keep; ++k) dst[k] = make_int2(X0[k * 64 + lane], X1[k * 64 + lane]);
    }
    reinterpret_cast<SpanCnt *>(B.shead + RR_COLD(shdr_cap))[it].cnt[lane] =
        fits ? static_cast<uint16_t>(keep) : static_cast<uint16_t>(0);
    hdr_head(B.shead + it, fits ? static_cast<long long>(off) : -1ll, fx0, fx1, cx0, cx1);
    // publish: composite tiles this opaque item covers fully (the same test
    // as the composite pass's occlusion scan)
    if (occlusion && Y0 + kBandRows <= P.height && fx0 <= fx1) {
        const int32_t ta = (fx0 + (1 << sh) - 1) >> sh;
        for (int32_t t = ta + lane; t < P.ctiles; t += 64) {
            const int32_t txe = min((t + 1) << sh, P.width) - 1;
            if (txe > fx1) break;
            atomicMax(const_cast<int32_t *>(occ_row) + t, static_cast<int32_t>(P.p));
        }
    }
}

// Header of an item skipped as occluded (its interval counts are never read);
// the _lane form is stored by the calling lane.
__device__ __forceinline__ void write_skipped(SpanHead *H) { hdr_head(H, kSkipped, 1, 0, 1, 0); }
__device__ __forceinline__ void hdr_skipped_lane(SpanHead *H) {
    *reinterpret_cast<int4 *>(H) = make_int4(static_cast<int32_t>(kSkipped & 0xffffffffll),
                                             static_cast<int32_t>(kSkipped >> 32), 1, 1);
}

// One (path, band).  sm0: the path's first chunk summaries
// (chunk_summary_first).  check: test the band for occlusion first (else the
// caller has); occlusion: publish the tiles an opaque band covers.  DEFER: a
// band with more state-changing entries on a row than the K LDS slots hold
// goes to the deferred queue (the main pass); otherwise it is redone at once
// with the entry lists in global memory (the wide pass).
template <int K, bool DEFER>
__device__ __forceinline__ void span_band(const BatchArgs &B, SpanLds<K> &L, const ItemPath &P, int32_t Y0,
                                          long long it, int64_t slot, bool count_draw, bool occlusion, bool check,
                                          uint64_t vmask, int2 sm0, SpanCursor &cur) {
    const FillArgs &A = B.F;
    const int lane = threadIdx.x;
    const int32_t band = Y0 / kBandRows;
    const int32_t *occ_row = B.occ + P.occ_base + static_cast<long long>(band) * P.ctiles;
    if (check && P.px0 <= P.px1) {
        // Every composite tile this item can reach already has a later opaque
        // path covering it fully: the composite pass starts after that path,
        // so these intervals would never be read.  Should the path be widened
        // later by an unbalanced row, the fixup pass redoes the item.
        // vmask: the first 64 tiles; beyond them only whether any tile is
        // visible (the loop is wave-uniform: every lane takes the same branch
        // below, so the wave stays whole for the build and the pool
        // allocation of span_finish)
        const int32_t t0 = P.px0 >> B.chunk_shift, t1 = P.px1 >> B.chunk_shift;
        const bool seen = t0 + lane <= t1 && occ_row[t0 + lane] <= P.p;
        vmask = __ballot(seen);
        bool visible = vmask != 0;
        for (int32_t tb = t0 + 64; !visible && tb <= t1; tb += 64)
            visible = __ballot(tb + lane <= t1 && occ_row[tb + lane] <= P.p) != 0;
        if (!visible) {
            ++cur.skipped;
            write_skipped(B.shead + it);
            return;
        }
    }
    // the path has a band that is built (pstate bit 1; for the statistics)
    if (P.p != cur.marked) {
        if (lane == 0) atomicOr(RR_COLD(pstate) + P.p, 2);
        cur.marked = static_cast<int32_t>(P.p);
    }
    const bool literal = RR_COLD(F.mode) == RR_MODE_LITERAL;
    const int32_t y = Y0 + (lane >> 2), j = lane & 3;
    const int32_t r = literal ? 16 * y + 12 + j : 4 * y + j;
    const int32_t band_lo = literal ? 16 * Y0 + 12 : 4 * Y0;
    const int32_t band_hi = literal ? 16 * (Y0 + kBandRows - 1) + 15 : 4 * Y0 + 4 * kBandRows - 1;
    const bool valid = y >= P.row0 && y <= P.row1 && r >= P.y_min && r <= P.y_max;
    const int32_t mrule = P.eo ? 1 : -1;
    int n = build_entries(A, L, P.e0, P.e1, band_lo, band_hi, literal, Y0, valid, r, mrule, L.ex0, L.ex1, L.ec, K,
                          sm0);
    if (__ballot(n > K) == 0) {
        bool unb = false;
        if (__ballot(n > 2) == 0) {
            RegIval s0{0}, e0{0};
            const int nint = walk_short(n, L.ex0, L.ex1, L.ec, mrule, P.width * 4, &unb, &s0.v, &e0.v);
            span_finish<1>(B, P, Y0, it, occ_row, s0, e0, nint, unb, count_draw, occlusion, vmask, cur);
        } else {
            const int nint = sort_and_walk(n, L.ex0, L.ex1, L.ec, mrule, P.width * 4, &unb);
            int32_t *X0 = L.ex0, *X1 = L.ex1;
            span_finish<0>(B, P, Y0, it, occ_row, X0, X1, nint, unb, count_draw, occlusion, vmask, cur);
        }
        return;
    }
    if constexpr (DEFER) {
        if (lane == 0) {
            const unsigned long long q = atomicAdd(RR_COLD(counters) + kCntDeferred, 1ull);
            if (static_cast<long long>(q) < RR_COLD(ovq_cap))
                RR_COLD(ovq)[q] = make_int4(static_cast<int32_t>(P.p), Y0, static_cast<int32_t>(it), 0);
            else
                atomicOr(RR_COLD(F.err), 32);
        }
        return;
    }
    // more state-changing entries on a row than the LDS holds: redo the band
    // with the entry lists in this wave's global slot or the shared pool
    const int32_t need = wave_max(n);
    int32_t *G;
    if (static_cast<long long>(need) * 192 <= A.slot_ints) {
        G = A.slots + slot * A.slot_ints;
    } else {
        unsigned long long off = 0;
        if (lane == 0) off = atomicAdd(A.spill_top, static_cast<unsigned long long>(need) * 192ull);
        off = first_lane_u64(off);
        if (static_cast<long long>(off) + static_cast<long long>(need) * 192 > A.spill_cap) {
            if (lane == 0) atomicOr(RR_COLD(F.err), 1);
            hdr_head(B.shead + it, -1, 1, 0, 1, 0);
            span_cnt(B, it)->cnt[lane] = 0;
            return;
        }
        G = A.spill + off;
    }
    int32_t *G1 = G + static_cast<long long>(need) * 64, *G2 = G + static_cast<long long>(need) * 128;
    n = build_entries(A, L, P.e0, P.e1, band_lo, band_hi, literal, Y0, valid, r, mrule, G, G1, G2, need, sm0);
    bool unb = false;
    const int nint = sort_and_walk(n, G, G1, G2, mrule, P.width * 4, &unb);
    span_finish<0>(B, P, Y0, it, occ_row, G, G1, nint, unb, count_draw, occlusion, vmask, cur);
}

constexpr int kSpanGroup = 16;   // items tested for occlusion together (4 lanes each)
constexpr int kGroupTiles = 16;  // composite tiles of an item tested there (4 per lane)

// Work queue split into kSlices contiguous slices with one counter each (in
// its own 256-byte line).  A wave starts on slice blockIdx.x % kSlices --
// workgroups are dispatched to the 8 XCDs round-robin, so each XCD works
// through its own slice, out of its own L2 -- and moves on to the next slice
// when its own is used up.  Each grab takes `grab` items; returns the first
// item (n when every slice is used up) and sets *end.
constexpr int kSlices = 8;
constexpr int kSliceStride = 32;  // counters per slice line (unsigned long long)
struct SliceCursor {
    int32_t slice, tried;
};
__device__ __forceinline__ long long slice_grab(unsigned long long *q, long long n, long long grab, SliceCursor &c,
                                                long long *end) {
    while (c.tried < kSlices) {
        const long long lo = n * c.slice / kSlices, hi = n * (c.slice + 1) / kSlices;
        unsigned long long g = 0;
        if (threadIdx.x == 0) g = atomicAdd(q + c.slice * kSliceStride, static_cast<unsigned long long>(grab));
        const long long g0 = lo + static_cast<long long>(first_lane_u64(g));
        if (g0 < hi) {
            *end = g0 + grab < hi ? g0 + grab : hi;
            return g0;
        }
        ++c.tried;
        c.slice = (c.slice + 1) % kSlices;
    }
    *end = n;
    return n;
}

constexpr int kSpanGrab = 64;  // span items per grab at most (fewer when the batch is small)
constexpr int kWideK = 32;     // LDS entry slots of the wide pass

// The path record of item path p as wave-uniform values (scalar loads).
__device__ __forceinline__ ItemPath item_path(const BatchArgs &B, int32_t p) {
    // through the constant address space: scalar loads (the records are
    // written by an earlier kernel and only read here)
    const RR_KARG_AS SpanPath *S = ((const RR_KARG_AS SpanPath *)(B.spath)) + p;
    ItemPath P;
    P.p = p;
    P.e0 = S->e0;
    P.e1 = S->e1;
    P.y_min = S->y_min;
    P.y_max = S->y_max;
    P.row0 = S->row0;
    P.row1 = S->row1;
    P.px0 = S->px0;
    P.px1 = S->px1;
    P.width = S->width;
    P.height = S->height;
    P.px_hi = S->px_hi;
    P.occ_base = S->occ_base;
    P.ctiles = S->ctiles;
    const uint32_t fl = S->flags;
    P.eo = fl & 1;
    P.opaque = (fl >> 1) & 1;
    P.px_lo = (fl >> 16) & 0xffff;
    return P;
}

// Main span pass: every work item, descending path order (occluders are met
// first), entry lists in LDS only; bands whose rows need more than K entries
// are deferred to the wide pass.  Items are taken kSpanGrab at a time from
// the sliced queue and handled kSpanGroup at a time; the first kGroupTiles
// composite tiles of the group's single-band items are tested for occlusion
// together (4 lanes per item, 4 tiles per lane), so a run of hidden items
// costs one round trip (a wider item none of whose first tiles is seen gets
// the full test of span_band), and a path's chunk summaries are loaded once
// for its run of consecutive items.  TRACE: record one timing record per item (a separate
// instance, so that the untraced pass carries none of its state).
template <int K, bool TRACE>
__global__ void __launch_bounds__(64) k_spans(BatchArgs B, int32_t mode) {
    __shared__ SpanLds<K> L;
    const FillArgs &A = B.F;
    const int lane = threadIdx.x;
    const long long n_items = min(B.work_base[B.n_paths], B.sitems_cap);
    const bool count_draw = mode == SPAN_ALL;
    const bool occlusion = (mode == SPAN_ALL || mode == SPAN_ALL_NODRAW) && !(A.cov_hash || A.cov_count);
    SpanCursor cur;
    int32_t sm_path = -1;
    int2 sm0 = make_int2(INT32_MAX, INT32_MIN);
    SliceCursor sc{static_cast<int32_t>(blockIdx.x % kSlices), 0};
    // about eight grabs per wave, so that a small batch still spreads over
    // the whole grid
    const long long grab = max(1ll, min(static_cast<long long>(kSpanGrab), n_items / (8ll * gridDim.x)));
    for (;;) {
        long long g_end = 0;
        const long long g0 = slice_grab(RR_COLD(queues), n_items, grab, sc, &g_end);
        if (g0 >= n_items) break;
      for (long long g = g0; g < g_end; g += kSpanGroup) {
        const int cnt = static_cast<int>(min(static_cast<long long>(kSpanGroup), g_end - g));
        // lanes 4k..4k+3 hold item k of the group; lanes 4k+2, 4k+3 quarters
        // 2 and 3 of its path record: {px0, px1, width, height},
        // {px_hi, occ_base, ctiles, flags}
        const int qi = lane >> 2, sub = lane & 3;
        int4 gi = make_int4(0, 0, 0, -1), gq = make_int4(0, 0, 0, 0);
        if (qi < cnt) {
            gi = RR_COLD(sitems)[n_items - 1 - g - qi];
            if (sub >= 2) gq = reinterpret_cast<const int4 *>(RR_COLD(spath) + gi.x)[sub];
        }
        const int32_t px0 = quad_bcast<2>(gq.x), px1 = quad_bcast<2>(gq.y);
        const int32_t obase = quad_bcast<3>(gq.y), ctl = quad_bcast<3>(gq.z);
        const bool live = qi < cnt && gi.w >= 0 && static_cast<long long>(gi.w) + gi.z <= RR_COLD(shdr_cap);
        if (qi < cnt && !live && sub == 0) atomicOr(RR_COLD(F.err), 2);
        // single-band items: lane sub tests tiles t0 + sub + 4i (i < 4), so the
        // first kGroupTiles tiles of each item are tested at once
        bool tested = false, wide = false, seen[4] = {false, false, false, false};
        if (occlusion && live && gi.z == 1 && px0 <= px1) {
            const int32_t t0 = px0 >> B.chunk_shift, t1 = px1 >> B.chunk_shift;
            const int32_t *row = B.occ + obase + static_cast<long long>(gi.y / kBandRows) * ctl;
            tested = true;
            wide = t1 - t0 >= kGroupTiles;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int32_t t = t0 + sub + 4 * i;
                if (t <= t1) seen[i] = row[t] <= gi.x;
            }
        }
        // an item none of whose tiles is seen is skipped here: its lane 4k
        // writes the header
        int any = (seen[0] || seen[1] || seen[2] || seen[3]) ? 1 : 0;
        any |= dpp_i32<kDppQuadXor1>(any);
        any |= dpp_i32<kDppQuadXor2>(any);
        const bool hidden = tested && !wide && any == 0;
        if (hidden && sub == 0) hdr_skipped_lane(B.shead + gi.w);
        cur.skipped += static_cast<uint32_t>(__popcll(__ballot(hidden && sub == 0)));
        const uint64_t wmask = __ballot(wide), tmask = __ballot(tested);
        const uint64_t smask0 = __ballot(seen[0]), smask1 = __ballot(seen[1]);
        const uint64_t smask2 = __ballot(seen[2]), smask3 = __ballot(seen[3]);
        // the group's other items, in order
        for (uint64_t todo = __ballot(live && !hidden && sub == 0); todo; todo &= todo - 1) {
            const int k = (__ffsll(static_cast<unsigned long long>(todo)) - 1) >> 2;
            const int32_t ip = __builtin_amdgcn_readlane(gi.x, 4 * k), iy = __builtin_amdgcn_readlane(gi.y, 4 * k);
            const int32_t iz = __builtin_amdgcn_readlane(gi.z, 4 * k), iw = __builtin_amdgcn_readlane(gi.w, 4 * k);
            // bit i of vm: tile t0 + i seen (not covered by a later opaque path)
            const bool was_tested = (tmask >> (4 * k)) & 1, was_wide = (wmask >> (4 * k)) & 1;
            const uint64_t vm = ((smask0 >> (4 * k)) & 0xf) | (((smask1 >> (4 * k)) & 0xf) << 4) |
                                (((smask2 >> (4 * k)) & 0xf) << 8) | (((smask3 >> (4 * k)) & 0xf) << 12);
            // a wide item none of whose first tiles is seen gets the full test
            const bool full_check = occlusion && (!was_tested || (was_wide && vm == 0));
            const uint64_t vmask = !was_tested ? ~0ull : (was_wide ? vm | (~0ull << kGroupTiles) : vm);
            const ItemPath P = item_path(B, ip);
            if (ip != sm_path) {
                sm0 = chunk_summary_first(A, P.e0, P.e1);
                sm_path = ip;
            }
            long long t0 = 0;
            if constexpr (TRACE) t0 = static_cast<long long>(wall_clock64());
            for (int32_t bnd = 0; bnd < iz; ++bnd)
                span_band<K, true>(B, L, P, iy + kBandRows * bnd, static_cast<long long>(iw) + bnd, 0, count_draw,
                                   occlusion, full_check, vmask, sm0, cur);
            if (TRACE && lane == 0) {
                const long long t1 = static_cast<long long>(wall_clock64());
                const unsigned long long kk = atomicAdd(A.trace_top, 1ull);
                if (static_cast<long long>(kk) < A.trace_cap)
                    A.trace[kk] = make_longlong4(ip, (static_cast<long long>(A.hdrs[ip].page) << 32) |
                                                         static_cast<uint32_t>(iy), t0, t1);
            }
        }
      }
    }
    if (lane == 0 && cur.stored) atomicAdd(RR_COLD(counters) + kCntStored, static_cast<unsigned long long>(cur.stored));
    if (lane == 0 && cur.skipped) atomicAdd(RR_COLD(counters) + kCntSkipped, static_cast<unsigned long long>(cur.skipped));
}

// Wide span pass: the bands the main pass deferred (mode = the main pass's
// mode), or the items skipped as occluded of paths an unbalanced row widened
// afterwards (SPAN_FIXUP).  kWideK LDS slots, then the wave's global slot.
template <int K>
__global__ void __launch_bounds__(64) k_spans_wide(BatchArgs B, int32_t mode) {
    __shared__ SpanLds<K> L;
    const int lane = threadIdx.x;
    const bool fixup = mode == SPAN_FIXUP;
    const long long n = fixup ? min(B.work_base[B.n_paths], B.sitems_cap)
                              : min(static_cast<long long>(B.counters[kCntDeferred]), B.ovq_cap);
    if (fixup && B.counters[kCntWidened] == 0) return;
    const bool count_draw = mode == SPAN_ALL;
    const bool occlusion = (mode == SPAN_ALL || mode == SPAN_ALL_NODRAW) && !(B.F.cov_hash || B.F.cov_count);
    unsigned long long *const served = B.counters + (fixup ? kCntFixupServed : kCntWideServed);
    SpanCursor cur;
    for (;;) {
        unsigned long long k = 0;
        if (lane == 0) k = atomicAdd(served, 1ull);
        k = first_lane_u64(k);
        if (static_cast<long long>(k) >= n) break;
        if (fixup) {
            const int4 item = B.sitems[k];
            if (item.w < 0 || static_cast<long long>(item.w) + item.z > B.shdr_cap || (B.pstate[item.x] & 1) == 0)
                continue;
            const ItemPath P = item_path(B, item.x);
            for (int32_t bnd = 0; bnd < item.z; ++bnd) {
                const long long hi = static_cast<long long>(item.w) + bnd;
                if (B.shead[hi].off != kSkipped) continue;
                span_band<K, false>(B, L, P, item.y + kBandRows * bnd, hi, blockIdx.x, false, false, false, ~0ull,
                                    chunk_summary_first(B.F, P.e0, P.e1), cur);
            }
        } else {
            const int4 q = B.ovq[k];
            const ItemPath P = item_path(B, q.x);
            span_band<K, false>(B, L, P, q.y, q.z, blockIdx.x, count_draw, occlusion, occlusion, ~0ull,
                                chunk_summary_first(B.F, P.e0, P.e1), cur);
        }
    }
    if (lane == 0 && cur.stored) atomicAdd(RR_COLD(counters) + kCntStored, static_cast<unsigned long long>(cur.stored));
    if (lane == 0 && cur.skipped) atomicAdd(RR_COLD(counters) + kCntSkipped, static_cast<unsigned long long>(cur.skipped));
}

// ---------------------------------------------------------------------------
// phase B: composite of one (page, band, tile)
// ---------------------------------------------------------------------------
// The tile's canvas bytes live in LDS as NPL channel planes (1: a grey or
// grey-Rgb8 canvas; 3: an Rgb8 canvas, de-interleaved on load and
// re-interleaved on store), 16 rows of CH pixels each.
template <int CH, int NPL>
struct ComposeLds {
    static constexpr int kWords = CH / 8;  // bitmap dwords per AA row (4 bits per pixel)
    static constexpr int kPitch = kWords + 1;
    static constexpr int kRowWords = CH / 4;  // canvas dwords per output row and plane
    static constexpr int kRowPitch = kRowWords + 1;
    static constexpr int kPlane = kBandRows * kRowPitch;
    uint32_t cv[NPL * kPlane];
    uint32_t bm[64 * kPitch];
    uint8_t gamma[32];
};

// One path of a tile bin as the composite pass consumes it.
struct PathRef {
    int64_t p;
    int32_t hi;   // span item
    uint32_t at;  // rr_path_attr bytes {flags, color, a_input, route}
    uint32_t rgb;  // Rgb8 canvas: the path's colour r | g << 8 | b << 16
    int32_t cx0, cx1;  // columns its intervals touch (from the span head)
};

// Walks a tile bin in order, 64 entries per chunk.  A chunk already loaded
// (the occlusion search's last one) is taken over when the walk starts in it.
struct BinCursor {
    long long lb, l1;
    int32_t k, n;
    int4 ent;
};

__device__ __forceinline__ bool bin_next(const BatchArgs &B, const DevPage &pg, BinCursor &c, PathRef &out,
                                         const uint32_t *colors) {
    const int lane = threadIdx.x;
    if (c.k >= c.n) {
        c.lb += 64;
        if (c.lb >= c.l1) return false;
        c.n = static_cast<int32_t>(min(64ll, c.l1 - c.lb));
        c.k = 0;
        c.ent = make_int4(0, 0, 0, 0);
        if (lane < c.n) c.ent = B.bins[c.lb + lane];
    }
    const int k = c.k++;
    out.hi = __builtin_amdgcn_readlane(c.ent.x, k);
    out.p = __builtin_amdgcn_readlane(c.ent.y, k);
    out.at = static_cast<uint32_t>(__builtin_amdgcn_readlane(c.ent.w, k));
    if (out.p < pg.path_begin || out.p >= pg.path_end || out.hi < 0 || out.hi >= B.shdr_cap) {
        if (lane == 0) atomicOr(B.F.err, 16);  // a bin entry that was never written
        out.hi = 0;
        out.at = 0;
        out.p = pg.path_begin;
    }
    out.rgb = colors && out.p < pg.path_end ? colors[out.p] : 0u;
    return true;
}

// Stored spans of one path: the pool offset of this lane's intervals, their
// count, and the first interval (loaded ahead of use).
struct SpanView {
    long long base;
    int32_t n;
    int2 v0;
};

__device__ __forceinline__ void hdr_fetch(const BatchArgs &B, PathRef &P, long long *off, int32_t *n) {
    const int32_t h = __builtin_amdgcn_readfirstlane(P.hi);
    const int4 head = *reinterpret_cast<const int4 *>(B.shead + h);
    *off = static_cast<long long>((static_cast<uint64_t>(static_cast<uint32_t>(head.y)) << 32) |
                                  static_cast<uint32_t>(head.x));
    P.cx0 = static_cast<int16_t>(head.w & 0xffff);
    P.cx1 = static_cast<int16_t>(head.w >> 16);
    *n = span_cnt(B, h)->cnt[threadIdx.x];
}

__device__ __forceinline__ SpanView span_fetch(const BatchArgs &B, long long off, int32_t n) {
    SpanView v;
    const int32_t excl = wave_incl_add(n) - n;
    v.base = off + excl;
    v.n = off < 0 ? 0 : n;
    v.v0 = make_int2(INT32_MAX, INT32_MAX);
    if (v.n > 0) v.v0 = B.spans[v.base];
    return v;
}

// Path p over the tile held in LDS.  Lane l holds AA row l of the band; the
// four lanes of output row d = l/4 work that row together.  A row whose four
// AA rows each have at most one interval in the tile is done from the
// interval ends: pixels covered by all four (coverage 16) are written whole
// dwords at a time, the rest get their coverage count in closed form
// (sum over AA rows of the interval's overlap with [4x, 4x+4)).  Rows with
// more intervals go through the 4x4-bit coverage bitmap.  Both give the
// nibble-popcount coverage of fill/mod.rs:344-401 exactly.  Each covered
// pixel is painted in every channel plane with paint_px: the AA path, or the
// general pipe for a Multiply or soft-masked path (its soft-mask byte read
// from the page's mask).
// COV: per-path coverage outputs (hashes / counts) were asked for; the kernel
// instance without them carries none of that code.
template <int CH, bool COV, int NPL>
__device__ __forceinline__ void compose_rows(const BatchArgs &B, ComposeLds<CH, NPL> &L, const DevPage &pg, int64_t p,
                                             uint32_t at, uint32_t rgb, const SpanView &V, int32_t Y0, int32_t tx0,
                                             int32_t xa, int32_t xb, int32_t ya, int32_t yb) {
    const FillArgs &A = B.F;
    const int lane = threadIdx.x, d = lane >> 2, j = lane & 3, g0 = lane & ~3;
    using LdsT = ComposeLds<CH, NPL>;
    constexpr int kPitch = LdsT::kPitch, kRowPitch = LdsT::kRowPitch, kPlane = LdsT::kPlane;
    const uint32_t a_in = (at >> 16) & 0xffu, fl = at & 0xffu;
    const bool opaque = a_in == 255u && !(fl & kFillGeneral);
    const bool fast = A.flavour != RR_FLAVOUR_GRAY8;
    uint32_t srcc[NPL];
#pragma unroll
    for (int c = 0; c < NPL; ++c) srcc[c] = NPL == 1 ? (at >> 8) & 0xffu : (rgb >> (8 * c)) & 0xffu;
    constexpr bool want_cov = COV;
    uint64_t hsum = 0;
    unsigned long long w_cov = 0;
    const int32_t yy = Y0 + d;
    const bool rowv = yy >= ya && yy <= yb;
    // soft-mask byte of pixel x of this row (255 when the path has none)
    const uint8_t *mrow = ((fl & RR_FILL_SOFT_MASK) && pg.smask) ? pg.smask + static_cast<long long>(yy) * pg.smask_stride
                                                                   : nullptr;
    auto sm_at = [&](int32_t x) -> uint32_t { return mrow ? mrow[x] : 255u; };
    const int32_t Alo = 4 * xa, Ahi = 4 * (xb + 1);
    const int32_t n = rowv ? V.n : 0;
    int32_t s0 = 0, e0 = 0;
    if (n >= 1) {
        s0 = max(V.v0.x, Alo);
        e0 = min(V.v0.y, Ahi);
        if (e0 <= s0) s0 = e0 = 0;
    }
    const int ne = e0 > s0 ? 1 : 0;
    int cx = n >= 2 ? 1 : 0;
    cx |= dpp_i32<kDppQuadXor1>(cx);
    cx |= dpp_i32<kDppQuadXor2>(cx);
    int all_ne = ne & dpp_i32<kDppQuadXor1>(ne);
    all_ne &= dpp_i32<kDppQuadXor2>(all_ne);
    int32_t fa = ne ? (s0 + 3) >> 2 : INT32_MIN, fb = ne ? (e0 >> 2) - 1 : INT32_MAX;
    int32_t ua = ne ? s0 >> 2 : INT32_MAX, ub = ne ? (e0 - 1) >> 2 : INT32_MIN;
    fa = max(fa, dpp_i32<kDppQuadXor1>(fa));
    fb = min(fb, dpp_i32<kDppQuadXor1>(fb));
    ua = min(ua, dpp_i32<kDppQuadXor1>(ua));
    ub = max(ub, dpp_i32<kDppQuadXor1>(ub));
    fa = max(fa, dpp_i32<kDppQuadXor2>(fa));
    fb = min(fb, dpp_i32<kDppQuadXor2>(fb));
    ua = min(ua, dpp_i32<kDppQuadXor2>(ua));
    ub = max(ub, dpp_i32<kDppQuadXor2>(ub));
    const int32_t sk[4] = {quad_bcast<0>(s0), quad_bcast<1>(s0), quad_bcast<2>(s0), quad_bcast<3>(s0)};
    const int32_t ek[4] = {quad_bcast<0>(e0), quad_bcast<1>(e0), quad_bcast<2>(e0), quad_bcast<3>(e0)};
    uint32_t *rowc = L.cv + d * kRowPitch;  // plane 0; plane c at + c * kPlane
    const int32_t dwbase = tx0 >> 2;
    const long long pixrow = static_cast<long long>(yy) * pg.width;
    if (rowv && !cx && ua <= ub) {
        int32_t dA = 1, dB = 0;  // whole dwords inside the fully covered run
        if (all_ne && fa <= fb) {
            dA = (fa + 3) >> 2;
            dB = ((fb + 1) >> 2) - 1;
        }
        for (int32_t w = dA + j; w <= dB; w += 4) {
#pragma unroll
            for (int c = 0; c < NPL; ++c) {
                uint32_t *cell = rowc + c * kPlane + (w - dwbase);
                uint32_t out = srcc[c] * 0x01010101u;
                if (!opaque) {
                    const uint32_t old = *cell;
                    out = 0;
#pragma unroll
                    for (int px = 0; px < 4; ++px)
                        out |= paint_px((old >> (8 * px)) & 0xffu, 255u, a_in, srcc[c], fl, sm_at(4 * w + px), fast)
                               << (8 * px);
                }
                *cell = out;
            }
            if (want_cov) {
                w_cov += 4;
                if (A.cov_hash)
                    for (int px = 0; px < 4; ++px) hsum += cov_term(pixrow + 4 * w + px, 255u);
            }
        }
        int32_t r1a = ua, r1b = ub, r2a = 1, r2b = 0;
        if (dA <= dB) {
            r1b = min(ub, 4 * dA - 1);
            r2a = max(ua, 4 * (dB + 1));
            r2b = ub;
        }
        const int32_t len1 = max(r1b - r1a + 1, 0), len2 = max(r2b - r2a + 1, 0);
        for (int32_t k = j; k < len1 + len2; k += 4) {
            const int32_t x = k < len1 ? r1a + k : r2a + (k - len1);
            const int32_t X0 = 4 * x, X1 = X0 + 4;
            int32_t cnt = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) cnt += max(0, min(ek[q], X1) - max(sk[q], X0));
            if (cnt == 0) continue;
            const uint32_t shape = L.gamma[cnt];
            const uint32_t sm = sm_at(x);
#pragma unroll
            for (int c = 0; c < NPL; ++c) {
                uint8_t *cb = reinterpret_cast<uint8_t *>(rowc + c * kPlane) + (x - tx0);
                const uint32_t res = (cnt == 16 && opaque) ? srcc[c] : paint_px(*cb, shape, a_in, srcc[c], fl, sm, fast);
                *cb = static_cast<uint8_t>(res);
            }
            if (want_cov) {
                ++w_cov;
                if (A.cov_hash) hsum += cov_term(pixrow + x, shape);
            }
        }
    } else if (rowv && cx) {
        // bitmap route: each lane sets its AA row's bits over [xa, xb]
        uint32_t *brow = L.bm + lane * kPitch;
        const int wa = (xa - tx0) >> 3, wb = (xb - tx0) >> 3;
        for (int w = wa; w <= wb; ++w) brow[w] = 0u;
        const int32_t A0 = 4 * tx0;
        for (int k = 0; k < n; ++k) {
            const int2 v = k == 0 ? V.v0 : B.spans[V.base + k];
            if (v.x >= Ahi) break;
            if (v.y <= Alo) continue;
            const int32_t lo = max(v.x, Alo) - A0, hi2 = min(v.y, Ahi) - A0;
            const int wl = lo >> 5, wh = (hi2 - 1) >> 5;
            const uint32_t ml = ~0u << (lo & 31);
            const uint32_t mh = (hi2 & 31) ? ((1u << (hi2 & 31)) - 1u) : ~0u;
            if (wl == wh) {
                brow[wl] |= ml & mh;
            } else {
                brow[wl] |= ml;
                for (int w = wl + 1; w < wh; ++w) brow[w] = ~0u;
                brow[wh] |= mh;
            }
        }
        __builtin_amdgcn_wave_barrier();
        const int32_t dw0 = (xa - tx0) >> 2, dw1 = (xb - tx0) >> 2;
        for (int32_t w = dw0 + j; w <= dw1; w += 4) {
            const int sh = (w & 1) * 16;
            uint32_t cvv = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t f = (L.bm[(g0 + q) * kPitch + (w >> 1)] >> sh) & 0xffffu;
                uint32_t t = f - ((f >> 1) & 0x5555u);
                t = (t & 0x3333u) + ((t >> 2) & 0x3333u);
                cvv += (t & 0xfu) | ((t & 0xf0u) << 4) | ((t & 0xf00u) << 8) | ((t & 0xf000u) << 12);
            }
            if (cvv == 0) continue;
            const int32_t x0 = tx0 + 4 * w;
#pragma unroll
            for (int c = 0; c < NPL; ++c) {
                uint32_t *cell = rowc + c * kPlane + w;
                const uint32_t old = *cell;
                uint32_t out;
                if (cvv == 0x10101010u && opaque) {
                    out = srcc[c] * 0x01010101u;
                } else {
                    out = old;
#pragma unroll
                    for (int px = 0; px < 4; ++px) {
                        const uint32_t cnt = (cvv >> (8 * px)) & 0xffu;
                        if (cnt == 0) continue;
                        const uint32_t res =
                            paint_px((out >> (8 * px)) & 0xffu, L.gamma[cnt], a_in, srcc[c], fl, sm_at(x0 + px), fast);
                        out = (out & ~(0xffu << (8 * px))) | (res << (8 * px));
                    }
                }
                if (out != old) *cell = out;
            }
            if (want_cov) {
                const long long pix0 = pixrow + x0;
#pragma unroll
                for (int px = 0; px < 4; ++px) {
                    const uint32_t cnt = (cvv >> (8 * px)) & 0xffu;
                    if (cnt == 0) continue;
                    ++w_cov;
                    if (A.cov_hash) hsum += cov_term(pix0 + px, L.gamma[cnt]);
                }
            }
        }
    }
    if (want_cov) {
        for (int o = 32; o > 0; o >>= 1) {
            hsum += __shfl_xor(hsum, o);
            w_cov += __shfl_xor(w_cov, o);
        }
        if (lane == 0) {
            if (A.cov_hash && hsum) atomicAdd(reinterpret_cast<unsigned long long *>(A.cov_hash + p), hsum);
            if (A.cov_count && w_cov) atomicAdd(reinterpret_cast<unsigned long long *>(A.cov_count + 2 * p), w_cov);
        }
    }
}

// The Gray8 output of an Rgb8 tile read straight from the canvas (a tile no
// path touches keeps the canvas as it was; the epilogue still converts it).
__device__ __forceinline__ void tile_gray_direct(const DevPage &pg, int32_t Y0, int32_t tx0, int32_t nrows,
                                                 int32_t ndw) {
    for (int idx = threadIdx.x; idx < nrows * ndw; idx += 64) {
        const int r = idx / ndw, w = idx - (idx / ndw) * ndw;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(pg.canvas + static_cast<long long>(Y0 + r) * pg.stride +
                                                                 3 * (tx0 + 4 * w));
        uint32_t cr, cg, cb;
        deinterleave4(src[0], src[1], src[2], cr, cg, cb);
        *reinterpret_cast<uint32_t *>(pg.gray + static_cast<long long>(Y0 + r) * pg.gray_stride + tx0 + 4 * w) =
            gray4(cr, cg, cb);
    }
}

// The paths of tile bin `ci` (band `band`, tile `tile` of page `page`), in
// path order, composited into the tile held in LDS.  Two-deep software
// pipeline: while path j composites, the first span of path j+1 and the
// header of path j+2 are in flight.  On an Rgb8 page with a Gray8 output the
// tile's rgb_to_gray is written as it is stored (every tile of the page).
template <int CH, bool COV, int NPL>
__device__ void compose_tile(const BatchArgs &B, ComposeLds<CH, NPL> &L, long long ci, int32_t page, int32_t band,
                             int32_t tile, unsigned long long *n_paths, unsigned long long *n_tiles,
                             unsigned long long *n_filled, uint32_t *n_ivals) {
    const FillArgs &A = B.F;
    const int lane = threadIdx.x;
    using LdsT = ComposeLds<CH, NPL>;
    constexpr int kRowWords = LdsT::kRowWords, kRowPitch = LdsT::kRowPitch, kPlane = LdsT::kPlane;
    const DevPage &pg = A.pages[page];
    const int32_t Y0 = kBandRows * band, tx0 = CH * tile;
    const int32_t width = pg.width;
    const int32_t txe = min(tx0 + CH, width) - 1;
    const int32_t nrows = min(kBandRows, pg.height - Y0);
    const int32_t ndw = ((txe - tx0) >> 2) + 1;
    const bool gray = NPL == 3 && pg.gray != nullptr;
    const uint32_t *colors = NPL == 3 ? A.colors : nullptr;
    const long long b0 = B.bin_off[ci * kBinSegs], b1 = B.bin_off[(ci + 1) * kBinSegs];
    if (b1 <= b0 || txe < pg.px_lo || tx0 > pg.px_hi) {
        if (gray) tile_gray_direct(pg, Y0, tx0, nrows, ndw);
        return;
    }
    // Occlusion: the last opaque path that covers the whole tile on all 64 AA
    // rows leaves exactly its colour there (both composites give src for
    // coverage 16 at a_input 255), so compositing starts after it on a tile
    // filled with that colour.  The bin is searched from its end.  Skipped
    // when per-path coverage is counted.
    long long start = b0;
    bool filled = false;
    uint32_t fill_c = 0;  // the cover's colour, one byte per plane
    BinCursor cur;
    cur.lb = b0 - 64;
    cur.l1 = b1;
    cur.k = cur.n = 0;
    if (!COV && nrows == kBandRows) {
        long long last = -1, lb = b1 - 64;
        int4 e = make_int4(0, 0, 0, 0);
        for (;; lb -= 64) {
            const long long q = lb + lane;
            e = make_int4(0, 0, 0, 0);
            bool full = false;
            if (q >= b0 && q < b1) {
                e = B.bins[q];
                full = static_cast<int16_t>(e.z & 0xffff) <= tx0 && static_cast<int16_t>(e.z >> 16) >= txe;
            }
            const unsigned long long m = __ballot(full);
            if (m) {
                last = lb + 63 - __clzll(m);
                break;
            }
            if (lb <= b0) break;
        }
        if (last >= 0) {
            start = last + 1;
            filled = true;
            const int k = static_cast<int>(last - lb);
            if (NPL == 3) {
                const int32_t pc = __builtin_amdgcn_readlane(e.y, k);
                fill_c = (pc >= pg.path_begin && pc < pg.path_end) ? colors[pc] : 0u;
            } else {
                fill_c = (__builtin_amdgcn_readlane(e.w, k) >> 8) & 0xff;
            }
        }
        // the walk starts inside the chunk just searched
        cur.lb = lb;
        cur.n = static_cast<int32_t>(min(64ll, b1 - lb));
        cur.k = static_cast<int32_t>(start - lb);
        cur.ent = e;
    }
    PathRef P0, P1;
    const bool has0 = bin_next(B, pg, cur, P0, colors);
    if (!has0 && !filled) {
        if (gray) tile_gray_direct(pg, Y0, tx0, nrows, ndw);
        return;
    }
    ++*n_tiles;
    if (filled) ++*n_filled;
    uint8_t *const base = pg.canvas + static_cast<long long>(Y0) * pg.stride + NPL * tx0;
    const int32_t stride = pg.stride;
    long long off0 = -1, off1 = -1;
    int32_t n0 = 0, n1 = 0;
    bool has1 = false;
    if (has0) {
        hdr_fetch(B, P0, &off0, &n0);
        has1 = bin_next(B, pg, cur, P1, colors);
        if (has1) hdr_fetch(B, P1, &off1, &n1);
    }
    for (int idx = lane; idx < kBandRows * kRowWords; idx += 64) {
        const int r = idx / kRowWords, w = idx % kRowWords;
        if (r < nrows && w < ndw) {
            if (NPL == 1) {
                L.cv[r * kRowPitch + w] =
                    filled ? fill_c * 0x01010101u
                           : *reinterpret_cast<const uint32_t *>(base + static_cast<long long>(r) * stride + 4 * w);
            } else {
                uint32_t pl[3];
                if (filled) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) pl[c] = ((fill_c >> (8 * c)) & 0xffu) * 0x01010101u;
                } else {
                    const uint32_t *src = reinterpret_cast<const uint32_t *>(base + static_cast<long long>(r) * stride) + 3 * w;
                    deinterleave4(src[0], src[1], src[2], pl[0], pl[1], pl[2]);
                }
#pragma unroll
                for (int c = 0; c < NPL; ++c) L.cv[c * kPlane + r * kRowPitch + w] = pl[c];
            }
        }
    }
    SpanView V0;
    V0.n = 0;
    if (has0) V0 = span_fetch(B, off0, n0);
    __syncthreads();
    if (has0 && off0 < 0 && lane == 0) atomicOr(A.err, 8);
    for (; has0;) {
        // stored intervals lie within the path's drawn columns and rows
        // next path: its first span; the one after: its header
        SpanView V1;
        V1.n = 0;
        PathRef P2;
        bool has2 = false;
        long long off2 = -1;
        int32_t n2 = 0;
        if (has1) {
            V1 = span_fetch(B, off1, n1);
            if (off1 < 0 && lane == 0) atomicOr(A.err, 8);
            has2 = bin_next(B, pg, cur, P2, colors);
            if (has2) hdr_fetch(B, P2, &off2, &n2);
        }
        // only the columns the path's intervals touch (its span head's cx range)
        const int32_t xa = max(tx0, P0.cx0), xb = min(txe, P0.cx1);
        if (xa <= xb)
            compose_rows<CH, COV, NPL>(B, L, pg, P0.p, P0.at, P0.rgb, V0, Y0, tx0, xa, xb, Y0, Y0 + nrows - 1);
        ++*n_paths;
        *n_ivals += static_cast<uint32_t>(V0.n);
        __syncthreads();
        if (!has1) break;
        P0 = P1;
        V0 = V1;
        P1 = P2;
        has1 = has2;
        off1 = off2;
        n1 = n2;
    }
    for (int idx = lane; idx < kBandRows * kRowWords; idx += 64) {
        const int r = idx / kRowWords, w = idx % kRowWords;
        if (r < nrows && w < ndw) {
            if (NPL == 1) {
                *reinterpret_cast<uint32_t *>(base + static_cast<long long>(r) * stride + 4 * w) = L.cv[r * kRowPitch + w];
            } else {
                const uint32_t cr = L.cv[r * kRowPitch + w], cg = L.cv[kPlane + r * kRowPitch + w],
                               cb = L.cv[2 * kPlane + r * kRowPitch + w];
                uint32_t o0, o1, o2;
                interleave4(cr, cg, cb, o0, o1, o2);
                uint32_t *dst = reinterpret_cast<uint32_t *>(base + static_cast<long long>(r) * stride) + 3 * w;
                dst[0] = o0;
                dst[1] = o1;
                dst[2] = o2;
                if (gray)
                    *reinterpret_cast<uint32_t *>(pg.gray + static_cast<long long>(Y0 + r) * pg.gray_stride + tx0 +
                                                  4 * w) = gray4(cr, cg, cb);
            }
        }
    }
    __syncthreads();
}

template <int CH, bool COV, int NPL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5, 8))) k_compose(BatchArgs B, long long n_items) {
    __shared__ ComposeLds<CH, NPL> L;
    if (threadIdx.x < 17) L.gamma[threadIdx.x] = kGamma[threadIdx.x];
    __syncthreads();
    const int lane = threadIdx.x;
    unsigned long long n_paths = 0, n_tiles = 0, n_filled = 0;
    uint32_t n_ivals = 0;  // intervals this lane read
    // One item per grab: larger grabs unbalance the tail (heavy tiles come
    // in runs).
    SliceCursor sc{static_cast<int32_t>(blockIdx.x % kSlices), 0};
    long long it = 0, it_end = 0;
    // the page of the last item: consecutive items are mostly on one page
    CPage cp{0, 0, 0, 0};
    long long cp_lo = 1, cp_hi = 0;
    for (;;) {
        if (it >= it_end) {
            it = slice_grab(B.queues + kSlices * kSliceStride, n_items, 1, sc, &it_end);
            if (it >= n_items) break;
        }
        if (it < cp_lo || it >= cp_hi) {
            int32_t lo = 0, hi = B.n_cpages - 1;
            while (lo < hi) {
                const int32_t mid = (lo + hi + 1) / 2;
                if (static_cast<long long>(B.cpages[mid].base) <= it)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            cp = B.cpages[lo];
            cp_lo = cp.base;
            cp_hi = lo + 1 < B.n_cpages ? static_cast<long long>(B.cpages[lo + 1].base) : n_items;
        }
        const int32_t local = static_cast<int32_t>(it - cp.base);
        const int32_t band = local / cp.tiles, tile = local - band * cp.tiles;
        compose_tile<CH, COV, NPL>(B, L, it, cp.page, band, tile, &n_paths, &n_tiles, &n_filled, &n_ivals);
        ++it;
    }
    const int32_t ivals = __builtin_amdgcn_readlane(wave_incl_add(static_cast<int32_t>(n_ivals)), 63);
    if (lane == 0) {
        if (n_paths) atomicAdd(B.counters + kCntPathTiles, n_paths);
        if (n_tiles) atomicAdd(B.counters + kCntTiles, n_tiles);
        if (n_filled) atomicAdd(B.counters + kCntFilledTiles, n_filled);
        if (ivals) atomicAdd(B.counters + kCntIvalsRead, static_cast<unsigned long long>(static_cast<uint32_t>(ivals)));
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
#define RR_TRY(ctx, expr)                                                                                          \
    do {                                                                                                           \
        hipError_t e_ = (expr);                                                                                    \
        if (e_ != hipSuccess) return rrint::fail(ctx, RR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr long long kSlotInts = static_cast<long long>(kSlotK) * 192;

const void *span_kernel(int k, bool trace) {
    if (trace) {
        switch (k) {
            case 6: return reinterpret_cast<const void *>(&k_spans<6, true>);
            case 8: return reinterpret_cast<const void *>(&k_spans<8, true>);
            case 16: return reinterpret_cast<const void *>(&k_spans<16, true>);
            default: return reinterpret_cast<const void *>(&k_spans<32, true>);
        }
    }
    switch (k) {
        case 6: return reinterpret_cast<const void *>(&k_spans<6, false>);
        case 8: return reinterpret_cast<const void *>(&k_spans<8, false>);
        case 16: return reinterpret_cast<const void *>(&k_spans<16, false>);
        default: return reinterpret_cast<const void *>(&k_spans<32, false>);
    }
}

#define RR_COMPOSE(ch, cov, npl) reinterpret_cast<const void *>(&k_compose<ch, cov, npl>)
const void *compose_kernel(int ch, bool cov, int npl) {
    static const void *const table[2][2][4] = {
        {{RR_COMPOSE(64, false, 1), RR_COMPOSE(128, false, 1), RR_COMPOSE(256, false, 1), RR_COMPOSE(512, false, 1)},
         {RR_COMPOSE(64, true, 1), RR_COMPOSE(128, true, 1), RR_COMPOSE(256, true, 1), RR_COMPOSE(512, true, 1)}},
        {{RR_COMPOSE(64, false, 3), RR_COMPOSE(128, false, 3), RR_COMPOSE(256, false, 3), RR_COMPOSE(512, false, 3)},
         {RR_COMPOSE(64, true, 3), RR_COMPOSE(128, true, 3), RR_COMPOSE(256, true, 3), RR_COMPOSE(512, true, 3)}}};
    const int ci = ch == 64 ? 0 : (ch == 128 ? 1 : (ch == 256 ? 2 : 3));
    return table[npl == 3 ? 1 : 0][cov ? 1 : 0][ci];
}
#undef RR_COMPOSE

template <typename T>
int32_t fill(rr_hip_ctx *c, T *p, long long n, T value) {
    if (n <= 0) return RR_OK;
    const long long blocks = std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, c->stream, p, n, value);
    RR_TRY(c, hipGetLastError());
    return RR_OK;
}

// Exclusive scan of n inputs into out (may alias in when the types match);
// total (optional, device) receives the sum.  tmp holds ceil(n / 1024) values.
template <typename Tin, typename Tout>
int32_t scan_exclusive(rr_hip_ctx *c, const Tin *in, Tout *out, long long n, Tout *total, Tout *tmp) {
    if (n <= 0) {
        if (total) return fill<Tout>(c, total, 1, static_cast<Tout>(0));
        return RR_OK;
    }
    const long long nb = (n + kScanTile - 1) / kScanTile;
    if (nb > 0x7fffffffll) return rrint::fail(c, RR_E_INVARIANT, "scan: too many elements");
    hipLaunchKernelGGL((k_scan_reduce<Tin, Tout>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, c->stream,
                       in, n, tmp);
    RR_TRY(c, hipGetLastError());
    hipLaunchKernelGGL(k_scan_sums<Tout>, dim3(1), dim3(kScanThreads), 0, c->stream, tmp, nb, total);
    RR_TRY(c, hipGetLastError());
    hipLaunchKernelGGL((k_scan_down<Tin, Tout>), dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, c->stream,
                       in, out, n, static_cast<const Tout *>(tmp));
    RR_TRY(c, hipGetLastError());
    return RR_OK;
}

int32_t launch_profiled(rr_hip_ctx *c, const void *kernel, unsigned grid, void **args) {
    bool ok = true;
    hipEvent_t *ev = rrint::prof_pair(c, &ok);
    if (!ok) return rrint::fail(c, RR_E_DEVICE, "profiling event allocation failed");
    if (ev) RR_TRY(c, hipEventRecord(ev[0], c->stream));
    RR_TRY(c, hipLaunchKernel(kernel, dim3(grid), dim3(64), args, 0, c->stream));
    if (ev) RR_TRY(c, hipEventRecord(ev[1], c->stream));
    rrint::count_launch(c);
    return RR_OK;
}

template <typename T>
int32_t read_back(rr_hip_ctx *c, const T *dev, T *host) {
    RR_TRY(c, hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, c->stream));
    RR_TRY(c, hipStreamSynchronize(c->stream));
    return RR_OK;
}

}  // namespace

namespace rrint {

int32_t scan_counts(rr_hip_ctx *c, const int32_t *in, long long *out, long long n) {
    const int32_t rc = ensure(c, &c->scan_tmp, &c->scan_tmp_cap, static_cast<size_t>((n + kScanTile - 1) / kScanTile + 1));
    if (rc) return rc;
    return scan_exclusive<int32_t, long long>(c, in, out, n, out + n, c->scan_tmp);
}

int32_t init_batched(rr_hip_ctx *c) {
    for (int v = 0; v < kNumVariants; ++v) {
        int a = 0, b = 0;
        int b3 = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, span_kernel(kVariantK[v], false), 64, 0) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, compose_kernel(kVariantCH[v], false, 1), 64, 0) !=
                hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&b3, compose_kernel(kVariantCH[v], false, 3), 64, 0) !=
                hipSuccess)
            return fail(c, RR_E_DEVICE, "occupancy query failed");
        c->span_waves[v] = std::max(1, c->n_cu * std::max(1, a));
        c->compose_waves[v] = std::max(1, c->n_cu * std::max(1, b));
        c->compose_waves3[v] = std::max(1, c->n_cu * std::max(1, b3));
    }
    int w = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&w, reinterpret_cast<const void *>(&k_spans_wide<kWideK>), 64, 0) !=
        hipSuccess)
        return fail(c, RR_E_DEVICE, "occupancy query failed");
    c->wide_waves = std::max(1, c->n_cu * std::max(1, w));
    return RR_OK;
}

int32_t record_batched(rr_hip_ctx *c, const rr_fill_pages &b, const FillArgs &A, int k_entries, int chunk_px,
                       int variant) {
    const long long n_paths = b.n_paths;
    if (n_paths >= 0x7fffffffll) return fail(c, RR_E_INVARIANT, "batched launch: too many paths");
    long long n_bands = 0;
    for (int32_t i = 0; i < b.n_pages; ++i) n_bands += (b.pages[i].height + kBandRows - 1) / kBandRows;
    if (n_bands * kBinSegs >= 0x7fffffffll) return fail(c, RR_E_INVARIANT, "batched launch: too many bands");
    int32_t rc = ensure(c, &c->item_base, &c->item_base_cap, static_cast<size_t>(n_paths + 1));
    const long long scan_n = std::max<long long>(std::max<long long>(n_paths, n_bands), c->batch_citems * kBinSegs);
    if (!rc) rc = ensure(c, &c->scan_tmp, &c->scan_tmp_cap, static_cast<size_t>((scan_n + kScanTile - 1) / kScanTile + 1));
    if (!rc) rc = ensure(c, &c->band_cnt, &c->band_cnt_cap, static_cast<size_t>(n_bands + 2));
    if (!rc) rc = ensure(c, &c->band_run, &c->band_run_cap, static_cast<size_t>(n_bands + 2));
    if (!rc) rc = ensure(c, &c->band_base, &c->band_base_cap, static_cast<size_t>(n_bands + 1));
    if (!rc) rc = ensure(c, &c->cpages, &c->cpages_cap, static_cast<size_t>(b.n_pages));
    if (rc) return rc;

    long long n_citems = c->batch_citems;  // set with the page table by the caller
    rc = ensure(c, &c->occ, &c->occ_cap, static_cast<size_t>(std::max<long long>(n_citems, 1)));
    const long long n_binsegs = n_citems * kBinSegs;  // (bin, segment) counts
    if (!rc) rc = ensure(c, &c->bin_cnt, &c->bin_cnt_cap, static_cast<size_t>(std::max<long long>(n_binsegs, 1)));
    if (!rc) rc = ensure(c, &c->bin_off, &c->bin_off_cap, static_cast<size_t>(n_binsegs + 1));
    if (rc) return rc;
    if (!c->capturing)
        RR_TRY(c, hipMemcpyAsync(c->cpages, c->host_cpages.data(), sizeof(CPage) * b.n_pages,
                                 hipMemcpyHostToDevice, c->stream));

    // dispatch (classes + work items per path), then span items per path
    // (one per band) -> item_base, work items -> work_base
    rc = ensure(c, &c->route_cls, &c->route_cls_cap, static_cast<size_t>(n_paths));
    if (!rc) rc = ensure(c, &c->spath, &c->spath_cap, static_cast<size_t>(n_paths));
    if (!rc) rc = ensure(c, &c->n_work, &c->n_work_cap, static_cast<size_t>(n_paths));
    if (!rc) rc = ensure(c, &c->work_base, &c->work_base_cap, static_cast<size_t>(n_paths + 1));
    if (rc) return rc;
    hipLaunchKernelGGL(k_dispatch, dim3(static_cast<unsigned>((n_paths + 255) / 256)), dim3(256), 0, c->stream,
                       b.attrs, c->nb, c->sel, c->feat, c->mlp_loaded ? c->mlp : nullptr, n_paths, c->route_cls,
                       c->n_work, c->hdrs, c->pages, b.edge_off, c->spath);
    RR_TRY(c, hipGetLastError());
    rc = scan_exclusive<int32_t, long long>(c, c->nb, c->item_base, n_paths, c->item_base + n_paths, c->scan_tmp);
    if (!rc)
        rc = scan_exclusive<int32_t, long long>(c, c->n_work, c->work_base, n_paths, c->work_base + n_paths,
                                                c->scan_tmp);
    if (!rc) rc = fill<int32_t>(c, c->band_cnt, n_bands + 2, 0);
    if (!rc) rc = fill<unsigned long long>(c, c->counters, kCounters, 0ull);
    if (!rc) rc = fill<unsigned long long>(c, c->queues, kQueueWords, 0ull);
    if (!rc) rc = fill<unsigned long long>(c, c->spill_top, 1, 0ull);
    if (!rc) rc = fill<int32_t>(c, c->occ, n_citems, -1);
    if (!rc) rc = ensure(c, &c->pstate, &c->pstate_cap, static_cast<size_t>(n_paths));
    if (!rc) rc = fill<int32_t>(c, c->pstate, n_paths, 0);
    if (rc) return rc;
    long long n_items = c->batch_items, n_work = c->batch_work;  // while capturing: the sizing pass's counts
    if (!c->capturing) {
        rc = read_back(c, c->item_base + n_paths, &n_items);  // also lets the cpage copy land
        if (!rc) rc = read_back(c, c->work_base + n_paths, &n_work);
        if (rc) return rc;
    }
    if (n_items >= 0x7fffffffll) return fail(c, RR_E_INVARIANT, "batched launch: too many (path, band) items");
    const size_t cap_items = static_cast<size_t>(std::max<long long>(n_items, 1));
    rc = ensure(c, &c->sitems, &c->sitems_cap, static_cast<size_t>(std::max<long long>(n_work, 1)));
    if (!rc) rc = ensure(c, &c->shdr, &c->shdr_cap, cap_items);
    if (!rc) rc = ensure(c, &c->band_list, &c->band_list_cap, cap_items);
    if (!rc && c->spans_cap == 0) rc = ensure(c, &c->spans, &c->spans_cap, std::max<size_t>(cap_items * 24, 1u << 20));
    if (rc) return rc;
    if (n_items == 0) return RR_OK;

    hipLaunchKernelGGL(k_expand, dim3(static_cast<unsigned>((n_paths + 3) / 4)), dim3(256), 0, c->stream, c->sel,
                       c->nb, c->n_work, c->hdrs, c->pages, c->item_base, c->work_base, n_paths, c->sitems,
                       static_cast<long long>(c->sitems_cap), c->band_cnt, c->err_flag);
    RR_TRY(c, hipGetLastError());
    // paths per band = prefix sum of the marks (band_run[i + 1]); then the
    // bands' list offsets
    rc = scan_exclusive<int32_t, long long>(c, c->band_cnt, c->band_run, n_bands + 1, nullptr, c->scan_tmp);
    if (!rc) rc = scan_exclusive<long long, long long>(c, c->band_run + 1, c->band_base, n_bands,
                                                       c->band_base + n_bands, c->scan_tmp);
    if (rc) return rc;
    // the band lists are read only by the tile bins after the span pass, so
    // they are built on the auxiliary stream while the span pass runs (the
    // span pass's widening touches only the .w of the sel records they read);
    // with per-launch profiling on, in line, so that the span pass's events
    // time it alone
    const bool overlap = !c->profiling;
    if (overlap && !c->aux) {
        RR_TRY(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        RR_TRY(c, hipEventCreateWithFlags(&c->aux_fork, hipEventDisableTiming));
        RR_TRY(c, hipEventCreateWithFlags(&c->aux_join, hipEventDisableTiming));
    }
    if (overlap) {
        RR_TRY(c, hipEventRecord(c->aux_fork, c->stream));
        RR_TRY(c, hipStreamWaitEvent(c->aux, c->aux_fork, 0));
    }
    hipLaunchKernelGGL(k_band_lists, dim3(static_cast<unsigned>(n_bands)), dim3(64), 0, overlap ? c->aux : c->stream,
                       c->sel, c->pages, b.n_pages, c->item_base, c->band_base, c->band_list,
                       static_cast<long long>(c->band_list_cap), c->err_flag);
    RR_TRY(c, hipGetLastError());
    if (overlap) RR_TRY(c, hipEventRecord(c->aux_join, c->aux));

    // phase A
    const unsigned a_waves = static_cast<unsigned>(std::min<long long>(c->span_waves[variant], std::max(n_work, 1ll)));
    const unsigned w_waves = static_cast<unsigned>(c->wide_waves);
    rc = ensure(c, &c->slots, &c->slots_cap, static_cast<size_t>(w_waves) * kSlotInts);
    if (!rc) rc = ensure(c, &c->ovq, &c->ovq_cap, static_cast<size_t>(n_items / 16 + 65536));
    if (rc) return rc;
    BatchArgs B{};
    B.F = A;
    B.F.slots = c->slots;
    B.spath = c->spath;
    B.sel = c->sel;
    B.item_base = c->item_base;
    B.n_paths = n_paths;
    B.sitems = c->sitems;
    B.sitems_cap = static_cast<long long>(c->sitems_cap);
    B.work_base = c->work_base;
    B.shdr_cap = static_cast<long long>(c->shdr_cap);
    B.shead = reinterpret_cast<SpanHead *>(c->shdr);
    B.spans = c->spans;
    B.spans_cap = static_cast<long long>(c->spans_cap);
    B.band_base = c->band_base;
    B.band_list = c->band_list;
    B.cpages = c->cpages;
    B.n_cpages = b.n_pages;
    B.counters = c->counters;
    B.occ = c->occ;
    B.chunk_px = chunk_px;
    B.chunk_shift = __builtin_ctz(static_cast<unsigned>(chunk_px));
    B.pstate = c->pstate;
    B.ovq = c->ovq;
    B.ovq_cap = static_cast<long long>(c->ovq_cap);
    B.bin_off = c->bin_off;
    B.bins = c->bins;
    B.queues = c->queues;
    // tile bins of the composite pass: sizes, then offsets (the sizes follow
    // the span pass: skipped items and items without intervals are left out)
    auto count_bins = [&]() -> int32_t {
        if (overlap) RR_TRY(c, hipStreamWaitEvent(c->stream, c->aux_join, 0));
        hipLaunchKernelGGL(k_tile_bins<false>, dim3(static_cast<unsigned>(n_bands * kBinSegs)), dim3(64), 0, c->stream,
                           c->pages, b.n_pages, c->band_base, c->band_list, B.shead, static_cast<long long>(c->shdr_cap),
                           chunk_px, c->occ, b.attrs, c->bin_cnt, c->bin_off, c->bins, static_cast<long long>(c->bins_cap),
                           c->err_flag);
        RR_TRY(c, hipGetLastError());
        return scan_exclusive<int32_t, long long>(c, c->bin_cnt, c->bin_off, n_binsegs, c->bin_off + n_binsegs,
                                                  c->scan_tmp);
    };
    auto place_bins = [&]() -> int32_t {
        B.bins = c->bins;
        hipLaunchKernelGGL(k_tile_bins<true>, dim3(static_cast<unsigned>(n_bands * kBinSegs)), dim3(64), 0, c->stream,
                           c->pages, b.n_pages, c->band_base, c->band_list, B.shead, static_cast<long long>(c->shdr_cap),
                           chunk_px, c->occ, b.attrs, c->bin_cnt, c->bin_off, c->bins, static_cast<long long>(c->bins_cap),
                           c->err_flag);
        RR_TRY(c, hipGetLastError());
        return RR_OK;
    };
    // the sizing pass before a capture counts every interval (no skipping)
    int32_t mode = c->size_only ? SPAN_ALL_NOSKIP : SPAN_ALL;
    void *a_args[] = {&B, &mode};
    // composite pass: channel planes of the canvas and its persistent grid
    const int npl = b.flavour == RR_FLAVOUR_RGB8 ? 3 : 1;
    const unsigned b_waves = static_cast<unsigned>(
        std::min<long long>(npl == 3 ? c->compose_waves3[variant] : c->compose_waves[variant], n_citems));
    c->batch_valid = true;
    c->batch_paths = n_paths;
    c->batch_items = n_items;
    c->batch_work = n_work;
    c->batch_reruns = 0;
    c->prof_a0 = c->prof_used;
    for (int attempt = 0;; ++attempt) {
        const int32_t pass_mode = mode;
        rc = launch_profiled(c, span_kernel(k_entries, A.trace != nullptr), a_waves, a_args);
        // the bands the main pass deferred (exits at once when none)
        if (!rc) rc = launch_profiled(c, reinterpret_cast<const void *>(&k_spans_wide<kWideK>), w_waves, a_args);
        if (rc) return rc;
        if (pass_mode != SPAN_ALL_NOSKIP) {
            // skipped items of paths widened afterwards (exits at once when none)
            mode = SPAN_FIXUP;
            rc = launch_profiled(c, reinterpret_cast<const void *>(&k_spans_wide<kWideK>), w_waves, a_args);
            if (rc) return rc;
        }
        c->prof_a1 = c->prof_used;
        rc = count_bins();
        if (rc) return rc;
        if (c->capturing) {
            // sized by the eager pass: the pool holds every interval plus each
            // wave's unused chunk tail, the bins every visible item
            rc = place_bins();
            if (rc) return rc;
            c->prof_b = c->prof_used;
            void *b_args[] = {&B, &n_citems};
            return launch_profiled(c, compose_kernel(chunk_px, A.cov_hash || A.cov_count, npl), b_waves, b_args);
        }
        unsigned long long k[kCounters] = {};
        long long n_bin = 0;
        RR_TRY(c, hipMemcpyAsync(k, c->counters, sizeof(k), hipMemcpyDeviceToHost, c->stream));
        RR_TRY(c, hipMemcpyAsync(&n_bin, c->bin_off + n_binsegs, sizeof(n_bin), hipMemcpyDeviceToHost, c->stream));
        RR_TRY(c, hipStreamSynchronize(c->stream));
        const unsigned long long used = k[kCntPoolTop], deferred = k[kCntDeferred];
        c->batch_bins = n_bin;
        const bool pool_ok = static_cast<long long>(used) <= B.spans_cap;
        const bool queue_ok = static_cast<long long>(deferred) <= B.ovq_cap;
        if (c->size_only && pool_ok && queue_ok) {
            // the sizing pass stored every interval and deferral: its counts
            // are the whole batch's
            c->prof_b = c->prof_used;
            return RR_OK;
        }
        if (pool_ok && queue_ok) break;
        if (attempt >= 4) return fail(c, RR_E_CAPACITY, "span workspace still too small after growing it");
        // the interval pool or the deferred queue was too small: grow it and
        // redo the span pass (the canvas is untouched so far; the drawn-width
        // counts are complete)
        if (!pool_ok)
            rc = ensure(c, &c->spans, &c->spans_cap,
                        std::max(static_cast<size_t>(used + used / 4), 2 * static_cast<size_t>(B.spans_cap)));
        if (!rc && !queue_ok)
            rc = ensure(c, &c->ovq, &c->ovq_cap,
                        std::max(static_cast<size_t>(deferred + deferred / 4), 2 * static_cast<size_t>(B.ovq_cap)));
        if (!rc) rc = rrint::clear_err_bits(c, 32);  // the lost deferrals (other bits stay)
        if (!rc) rc = fill<unsigned long long>(c, c->counters, kCounters, 0ull);
        if (!rc) rc = fill<unsigned long long>(c, c->queues, kQueueWords, 0ull);
        if (!rc) rc = fill<unsigned long long>(c, c->spill_top, 1, 0ull);
        if (!rc) rc = fill<int32_t>(c, c->occ, n_citems, -1);
        if (!rc) rc = fill<int32_t>(c, c->pstate, n_paths, 0);
        if (rc) return rc;
        B.spans = c->spans;
        B.spans_cap = static_cast<long long>(c->spans_cap);
        B.ovq = c->ovq;
        B.ovq_cap = static_cast<long long>(c->ovq_cap);
        mode = c->size_only ? SPAN_ALL_NOSKIP : SPAN_ALL_NODRAW;
        c->batch_reruns = attempt + 1;
    }
    rc = ensure(c, &c->bins, &c->bins_cap, static_cast<size_t>(std::max<long long>(c->batch_bins, 1)));
    if (!rc) rc = place_bins();
    if (rc) return rc;
    c->prof_b = c->prof_used;

    // phase B
    void *b_args[] = {&B, &n_citems};
    return launch_profiled(c, compose_kernel(chunk_px, A.cov_hash || A.cov_count, npl), b_waves, b_args);
}

int32_t zero_u64(rr_hip_ctx *c, unsigned long long *p, long long n) { return fill<unsigned long long>(c, p, n, 0ull); }
int32_t fill_u64(rr_hip_ctx *c, unsigned long long *p, long long n, unsigned long long value) {
    return fill<unsigned long long>(c, p, n, value);
}
int32_t fill_i32(rr_hip_ctx *c, int32_t *p, long long n, int32_t value) { return fill<int32_t>(c, p, n, value); }
int32_t set_err_bits(rr_hip_ctx *c, int32_t bits) {
    hipLaunchKernelGGL(k_or_i32, dim3(1), dim3(1), 0, c->stream, c->err_flag, bits);
    RR_TRY(c, hipGetLastError());
    return RR_OK;
}
int32_t clear_err_bits(rr_hip_ctx *c, int32_t mask) {
    hipLaunchKernelGGL(k_and_i32, dim3(1), dim3(1), 0, c->stream, c->err_flag, ~mask);
    RR_TRY(c, hipGetLastError());
    return RR_OK;
}

}  // namespace rrint

struct rr_hip_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t ws_gen = 0;
    int64_t items = 0, citems = 0, work = 0;
    // the host-provided tables the sequence reads, owned by the graph so that
    // later recordings on the context cannot change what a replay sees
    DevPage *pages = nullptr;
    CPage *cpages = nullptr;
};

namespace {
void graph_free(rr_hip_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->pages) (void)hipFree(g->pages);
    if (g->cpages) (void)hipFree(g->cpages);
    delete g;
}
}  // namespace

extern "C" int32_t rr_hip_graph_create(rr_hip_ctx *c, const rr_fill_pages *b, rr_hip_graph **out) {
    if (!c || !b || !out) return RR_E_INVARIANT;
    *out = nullptr;
    if (c->page_open) return rrint::fail(c, RR_E_INVARIANT, "graph_create inside an open page");
    if (b->launch != RR_LAUNCH_BATCHED) return rrint::fail(c, RR_E_INVARIANT, "graph capture covers RR_LAUNCH_BATCHED");
    if (c->trace) return rrint::fail(c, RR_E_INVARIANT, "graph capture with a trace buffer set");
    RR_TRY(c, hipSetDevice(c->device));
    rr_fill_pages t = *b;
    const bool prof = c->profiling;
    c->profiling = false;
    // 1. eager sizing pass: tables uploaded, items counted, intervals counted
    //    (no composite, so the canvases are untouched)
    c->page_open = true;
    c->size_only = true;
    int32_t rc = rrint::record_pages_internal(c, t);
    c->size_only = false;
    c->page_open = false;
    unsigned long long k[kCounters] = {};
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = rrint::fail(c, RR_E_DEVICE, "sizing pass failed");
    if (!rc && hipMemcpy(k, c->counters, sizeof(k), hipMemcpyDeviceToHost) != hipSuccess)
        rc = rrint::fail(c, RR_E_DEVICE, "sizing pass read-back failed");
    if (!rc) {
        // every interval plus each wave's unused chunk tail (main and wide
        // passes); every band the main pass can defer (nothing is skipped in
        // the sizing pass, so replays defer no more)
        // (the sizing pass grew the pool and ran again until everything fit,
        // so the stored count is complete; the pool top bounds it too)
        const size_t waves = static_cast<size_t>(c->span_waves[t.variant]) + static_cast<size_t>(c->wide_waves);
        const size_t need = std::max(static_cast<size_t>(k[kCntStored]) + waves * static_cast<size_t>(kSpanChunk),
                                     static_cast<size_t>(k[kCntPoolTop])) + 1024;
        rc = rrint::ensure(c, &c->spans, &c->spans_cap, need);
        if (!rc) rc = rrint::ensure(c, &c->ovq, &c->ovq_cap, static_cast<size_t>(k[kCntDeferred]) + 1024);
        if (!rc) rc = rrint::ensure(c, &c->bins, &c->bins_cap, static_cast<size_t>(c->batch_bins) + 1024);
    }
    if (rc) {
        c->profiling = prof;
        return rc;
    }
    // 2. graph-owned copies of the page and composite-tile tables
    rr_hip_graph *g = new rr_hip_graph();
    g->items = c->batch_items;
    g->citems = c->batch_citems;
    g->work = c->batch_work;
    const size_t np = static_cast<size_t>(t.n_pages);
    if (hipMalloc(&g->pages, np * sizeof(DevPage)) != hipSuccess ||
        hipMalloc(&g->cpages, np * sizeof(CPage)) != hipSuccess ||
        hipMemcpy(g->pages, c->pages, np * sizeof(DevPage), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g->cpages, c->cpages, np * sizeof(CPage), hipMemcpyDeviceToDevice) != hipSuccess) {
        c->profiling = prof;
        graph_free(g);
        return rrint::fail(c, RR_E_DEVICE, "graph table copies failed");
    }
    // 3. capture the whole device sequence, reading the graph's tables
    DevPage *const ctx_pages = c->pages;
    CPage *const ctx_cpages = c->cpages;
    c->pages = g->pages;
    c->cpages = g->cpages;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        c->capturing = true;
        c->page_open = true;
        rc = rrint::record_pages_internal(c, t);
        c->page_open = false;
        c->capturing = false;
        const hipError_t e2 = hipStreamEndCapture(c->stream, &g->graph);
        if (!rc && e2 != hipSuccess) rc = rrint::fail(c, RR_E_DEVICE, std::string("end capture: ") + hipGetErrorString(e2));
    } else {
        rc = rrint::fail(c, RR_E_DEVICE, std::string("begin capture: ") + hipGetErrorString(e));
    }
    c->pages = ctx_pages;
    c->cpages = ctx_cpages;
    if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
        rc = rrint::fail(c, RR_E_DEVICE, "graph instantiate failed");
    c->profiling = prof;
    if (rc) {
        graph_free(g);
        return rc;
    }
    g->ws_gen = c->ws_gen;
    *out = g;
    return RR_OK;
}

extern "C" int32_t rr_hip_graph_launch(rr_hip_ctx *c, rr_hip_graph *g) {
    if (!c || !g) return RR_E_INVARIANT;
    if (!c->page_open) return rrint::fail(c, RR_E_INVARIANT, "graph_launch outside begin_page/submit_page");
    if (g->ws_gen != c->ws_gen)
        return rrint::fail(c, RR_E_INVARIANT, "the context's workspace was reallocated since capture: capture again");
    RR_TRY(c, hipSetDevice(c->device));
    bool ok = true;
    hipEvent_t *ev = rrint::prof_pair(c, &ok);
    if (!ok) return rrint::fail(c, RR_E_DEVICE, "profiling event allocation failed");
    if (ev) RR_TRY(c, hipEventRecord(ev[0], c->stream));
    RR_TRY(c, hipGraphLaunch(g->exec, c->stream));
    if (ev) RR_TRY(c, hipEventRecord(ev[1], c->stream));
    rrint::count_launch(c);
    c->batch_items = g->items;
    c->batch_citems = g->citems;
    c->batch_work = g->work;
    c->batch_reruns = 0;
    c->batch_valid = false;  // per-phase events are not inside the graph
    return RR_OK;
}

extern "C" int32_t rr_hip_graph_destroy(rr_hip_ctx *c, rr_hip_graph *g) {
    if (!g) return RR_E_INVARIANT;
    if (c) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    graph_free(g);
    return RR_OK;
}

namespace {
// Paths of the last batched recording with at least one band built (not
// skipped as occluded), and their edge records.
__global__ void __launch_bounds__(256) k_built_paths(const SpanPath *__restrict__ spath,
                                                     const int32_t *__restrict__ pstate, long long n_paths,
                                                     unsigned long long *__restrict__ out) {
    unsigned long long paths = 0, edges = 0;
    for (long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; p < n_paths;
         p += static_cast<long long>(gridDim.x) * blockDim.x) {
        if (pstate[p] & 2) {
            ++paths;
            edges += static_cast<unsigned long long>(spath[p].e1 - spath[p].e0);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        paths += __shfl_xor(paths, o);
        edges += __shfl_xor(edges, o);
    }
    if ((threadIdx.x & 63) == 0 && (paths || edges)) {
        atomicAdd(out, paths);
        atomicAdd(out + 1, edges);
    }
}
}  // namespace

extern "C" int32_t rr_hip_debug_batch_table(rr_hip_ctx *c, int32_t table, void *dst, int64_t cap, int64_t *count) {
    if (!c || !count || cap < 0 || (cap > 0 && !dst)) return RR_E_INVARIANT;
    if (!c->batch_valid && !c->batch_paths) return rrint::fail(c, RR_E_INVARIANT, "no batched recording");
    RR_TRY(c, hipSetDevice(c->device));
    RR_TRY(c, hipStreamSynchronize(c->stream));
    const void *src = nullptr;
    size_t elem = 0;
    long long n = 0;
    switch (table) {
        case 0: src = c->occ; elem = sizeof(int32_t); n = c->batch_citems; break;
        case 1: src = c->bin_off; elem = sizeof(long long); n = c->batch_citems * kBinSegs + 1; break;
        case 2: src = c->bins; elem = sizeof(int4); n = c->batch_bins; break;
        case 3: src = c->shdr; elem = sizeof(SpanHead); n = c->batch_items; break;
        case 4: src = c->item_base; elem = sizeof(long long); n = c->batch_paths + 1; break;
        case 5: src = c->spath; elem = sizeof(SpanPath); n = c->batch_paths; break;
        case 6: src = reinterpret_cast<const SpanHead *>(c->shdr) + c->shdr_cap; elem = sizeof(SpanCnt); n = c->batch_items; break;
        case 7: src = c->spans; elem = sizeof(int2); n = static_cast<long long>(c->spans_cap); break;
        default: return rrint::fail(c, RR_E_INVARIANT, "debug_batch_table: unknown table");
    }
    *count = n;
    const long long m = std::min<long long>(n, cap);
    if (m > 0 && src) RR_TRY(c, hipMemcpy(dst, src, static_cast<size_t>(m) * elem, hipMemcpyDeviceToHost));
    return RR_OK;
}

extern "C" int32_t rr_hip_last_batch_stats(rr_hip_ctx *c, double ms[3], int64_t counts[12]) {
    if (!c || !ms || !counts) return RR_E_INVARIANT;
    if (!c->batch_valid) return rrint::fail(c, RR_E_INVARIANT, "no batched recording since the context was created");
    RR_TRY(c, hipSetDevice(c->device));
    RR_TRY(c, hipStreamSynchronize(c->stream));
    unsigned long long k[kCounters] = {};
    RR_TRY(c, hipMemcpy(k, c->counters, sizeof(k), hipMemcpyDeviceToHost));
    ms[0] = ms[1] = ms[2] = 0.0;
    if (c->profiling && c->prof_b + 2 <= c->prof_used) {
        // the recording's span launches: the last attempt's main pass, then
        // its wide passes
        const size_t main = c->prof_a1 >= c->prof_a0 + 6 ? c->prof_a1 - 6 : c->prof_a0;
        for (size_t i = c->prof_a0; i + 1 < c->prof_a1; i += 2) {
            float m = 0.f;
            if (hipEventElapsedTime(&m, c->prof_ev[i], c->prof_ev[i + 1]) != hipSuccess) continue;
            if (i == main) ms[0] = m;
            else if (i > main) ms[1] += m;
        }
        float m = 0.f;
        if (hipEventElapsedTime(&m, c->prof_ev[c->prof_b], c->prof_ev[c->prof_b + 1]) == hipSuccess) ms[2] = m;
    }
    counts[0] = c->batch_items;
    counts[1] = static_cast<int64_t>(k[kCntStored]);
    counts[2] = c->batch_citems;
    counts[3] = static_cast<int64_t>(k[kCntPathTiles]);
    counts[4] = static_cast<int64_t>(k[kCntTiles]);
    counts[5] = c->batch_reruns;
    counts[6] = static_cast<int64_t>(k[kCntSkipped]);
    counts[7] = static_cast<int64_t>(k[kCntDeferred]);
    // paths with a built band and their edge records (a small reduction)
    unsigned long long built[2] = {0, 0};
    if (c->batch_paths > 0) {
        int32_t rc = fill<unsigned long long>(c, c->counters + kCntBuiltPaths, 2, 0ull);
        if (rc) return rc;
        const long long blocks = std::min<long long>((c->batch_paths + 255) / 256, 1024);
        hipLaunchKernelGGL(k_built_paths, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, c->stream, c->spath,
                           c->pstate, c->batch_paths, c->counters + kCntBuiltPaths);
        RR_TRY(c, hipGetLastError());
        RR_TRY(c, hipMemcpyAsync(built, c->counters + kCntBuiltPaths, sizeof(built), hipMemcpyDeviceToHost, c->stream));
        RR_TRY(c, hipStreamSynchronize(c->stream));
    }
    counts[8] = static_cast<int64_t>(built[0]);
    counts[9] = static_cast<int64_t>(built[1]);
    counts[10] = static_cast<int64_t>(k[kCntFilledTiles]);
    counts[11] = static_cast<int64_t>(k[kCntIvalsRead]);
    return RR_OK;
}

// Block sums of rr_hip_record_scan_static: a file-scope device array, so that
// the entry point allocates nothing.  Host code never uses the array's name as
// an address: each context keeps the per-device address hipGetSymbolAddress
// returned for it (rrint::init_static_scratch, at rr_hip_create).
constexpr uint32_t kStaticScanMax = 1u << 24;                   // elements per call at most
constexpr uint32_t kStaticScanSums = kStaticScanMax / kScanTile;  // 16384 uint32 (64 KiB)
static __device__ __align__(256) uint32_t scratch[kStaticScanSums];

namespace {
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[256] = {};  // per device ordinal
}  // namespace

namespace rrint {
int32_t init_static_scratch(rr_hip_ctx *c) {
    if (c->device < 0 || c->device >= 256) return fail(c, RR_E_INVARIANT, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // the device is current
        if (e != hipSuccess || !ptr)
            return fail(c, RR_E_DEVICE, std::string("static scratch address: ") + hipGetErrorString(e));
        g_scratch_ptr[c->device] = static_cast<uint32_t *>(ptr);
    }
    c->static_scratch = g_scratch_ptr[c->device];
    return RR_OK;
}
}  // namespace rrint

extern "C" int32_t rr_hip_record_scan_static(rr_hip_ctx *c, uint32_t *data, uint32_t len_elems) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (len_elems == 0) return rrint::fail(c, RR_E_INVARIANT, "scan: len_elems must be > 0");
    if (len_elems > kStaticScanMax)
        return rrint::fail(c, RR_E_INVARIANT, "scan_static: len_elems exceeds 16777216 (the static scratch holds "
                                              "16384 block sums); use rr_hip_record_scan");
    if (!data) return rrint::fail(c, RR_E_INVARIANT, "scan: null buffer");
    if (!c->static_scratch) return rrint::fail(c, RR_E_DEVICE, "static scratch not initialised");
    RR_TRY(c, hipSetDevice(c->device));
    return scan_exclusive<uint32_t, uint32_t>(c, data, data, len_elems, nullptr, c->static_scratch);
}

extern "C" int32_t rr_hip_record_scan(rr_hip_ctx *c, uint32_t *data, uint32_t *block_sums, uint32_t len_elems) {
    if (!c) return RR_E_INVARIANT;
    if (!rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "record outside begin_page/submit_page");
    if (len_elems == 0) return rrint::fail(c, RR_E_INVARIANT, "scan: len_elems must be > 0");
    if (!data || !block_sums) return rrint::fail(c, RR_E_INVARIANT, "scan: null buffer");
    RR_TRY(c, hipSetDevice(c->device));
    return scan_exclusive<uint32_t, uint32_t>(c, data, data, len_elems, nullptr, block_sums);
}
