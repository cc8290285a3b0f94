// rr_xpath.hip -- device path ingestion: XPath::new with flatten_curve on the
// GPU, bit-exact to the reference's f64 arithmetic (and to the host twin in
// rr_host.cpp), producing the add_segment edge list the fill kernels consume.
//
// XPath::new (crates/raster/src/xpath.rs:162-260, close_subpaths = true as
// fill_impl passes it, fill/mod.rs:125) walks a PathBuilder point list
// (path/mod.rs:70-79 flags) and emits, per subpath, a line per LINE point, the
// flattened pieces of each curve (flatten_curve, path/flatten.rs:109-215) and
// a closing line when the walk ends away from the subpath start.  Every
// segment starts at the point before it, so segments are independent once the
// walk has assigned each point its role:
//
//   k_xpath_walk     one wave per path: transform (xpath.rs:421-426) and
//                    role of every point, and where a closing line follows;
//   scan, compact    the points that emit edges (lines, curve starts, points a
//                    closing line follows), in order: the segments;
//   classify, order  segments grouped by the flatten work their curve needs;
//   k_xpath_count    one thread per segment: edges it emits (1 per line, the
//                    leaves of its curve, 1 per closing line), keeping the
//                    first 64 leaf points of a curve in a staging area;
//   scan             edge offset of every segment;
//   k_xpath_emit     one wave per 64 segments: writes their contiguous run of
//                    edges with coalesced stores, from the staged leaf points
//                    (a curve with more leaves is flattened again).
// Threads go to segments, not points, so the two control points that follow
// each curve start leave no lane idle.
//
// flatten_curve subdivides on a linked slot stack of 1024 slots: node [a, b)
// either is a leaf (b - a == 1, or both control points within flatness of the
// chord midpoint) and emits its end point, or splits at c = (a + b) / 2 and
// continues with [a, c) before [c, b): a depth-first walk of a binary tree at
// most ten levels deep.  The slots are dyadic intervals, so the node after a
// leaf ending at slot a is [a, a + lowbit(a)), and a node's control points are
// those obtained by splitting the root down to it: a thread keeps the current
// node and the three newest pending right halves in registers and rebuilds an
// older one from the root (at most ten splits), instead of the reference's
// 49 KB slot arrays.  Every value is produced by the same operations as in the
// reference.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "rr_hip.h"
#include "rr_internal.h"

namespace {

constexpr uint8_t kFirst = 1, kLast = 2, kCurve = 8;  // path/mod.rs:70-79
constexpr int32_t kSlots = 1024;                     // types.rs MAX_CURVE_SPLITS
constexpr uint8_t kRoleNone = 0, kRoleLine = 1, kRoleCurve = 2;

// f64::midpoint for finite operands: (a + b) / 2 unless that overflows or
// loses the low bits of a subnormal half.
constexpr double kHalfMax = 8.98846567431157953865e307;  // f64::MAX / 2

__host__ __device__ __forceinline__ double midpoint(double a, double b) {
    const double lo = 4.45014771701440227211e-308;  // 2 * f64::MIN_POSITIVE
    const double aa = fabs(a), ab = fabs(b);
    if (aa <= kHalfMax && ab <= kHalfMax) return (a + b) / 2.0;
    if (aa < lo) return a + b / 2.0;
    if (ab < lo) return a / 2.0 + b;
    return a / 2.0 + b / 2.0;
}

// midpoint() where both operands are known to lie in [-f64::MAX/2, f64::MAX/2]
// (HALF), where it is (a + b) / 2.
template <bool HALF>
__host__ __device__ __forceinline__ double mid(double a, double b) {
    return HALF ? (a + b) / 2.0 : midpoint(a, b);
}

struct Cubic {
    double x0, y0, x1, y1, x2, y2, x3, y3;
};

// Whether every coordinate of c lies in [-f64::MAX/2, f64::MAX/2] (false for
// NaN).  The midpoint of two values of magnitude at most M is one again
// (|a + b| <= 2M, and rounding to nearest cannot pass the representable
// bounds 2M and M), so the whole subdivision tree of such a root stays in
// range and every midpoint() in it takes its (a + b) / 2 branch.
__host__ __device__ __forceinline__ bool in_half_range(const Cubic &c) {
    return fabs(c.x0) <= kHalfMax && fabs(c.y0) <= kHalfMax && fabs(c.x1) <= kHalfMax && fabs(c.y1) <= kHalfMax &&
           fabs(c.x2) <= kHalfMax && fabs(c.y2) <= kHalfMax && fabs(c.x3) <= kHalfMax && fabs(c.y3) <= kHalfMax;
}

// De Casteljau halves of c (flatten.rs:186-205).
template <bool HALF>
__host__ __device__ __forceinline__ void split(const Cubic &c, Cubic &l, Cubic &r) {
    const double lx1 = mid<HALF>(c.x0, c.x1), ly1 = mid<HALF>(c.y0, c.y1);
    const double hx = mid<HALF>(c.x1, c.x2), hy = mid<HALF>(c.y1, c.y2);
    const double lx2 = mid<HALF>(lx1, hx), ly2 = mid<HALF>(ly1, hy);
    const double rx2 = mid<HALF>(c.x2, c.x3), ry2 = mid<HALF>(c.y2, c.y3);
    const double rx1 = mid<HALF>(hx, rx2), ry1 = mid<HALF>(hy, ry2);
    const double rx0 = mid<HALF>(lx2, rx1), ry0 = mid<HALF>(ly2, ry1);
    l = Cubic{c.x0, c.y0, lx1, ly1, lx2, ly2, rx0, ry0};
    r = Cubic{rx0, ry0, rx1, ry1, rx2, ry2, c.x3, c.y3};
}

// The leaves of flatten_curve in emission order: leaf(x, y) gets each piece's
// end point (the start of the first piece is the curve's p0).  Depth first,
// left half before right half.  The kKeep newest pending right halves are
// kept (without their first point, the end point of the last leaf emitted
// before them); an older one is rebuilt by splitting the root down to it.  A
// rebuild stalls every lane of the wave for its splits, so the kept depth is
// traded against registers and the moves of each push and pop.
struct Half {
    double x1, y1, x2, y2, x3, y3;
};
constexpr int kKeep = 3;

template <bool HALF, typename F>
__host__ __device__ __forceinline__ void flatten_leaves_t(const Cubic &root, double flat_sq, F &&leaf) {
    Cubic c = root;
    int32_t a = 0, size = kSlots;  // the current node is slots [a, a + size)
    int32_t pending = 0, kept = 0;  // pending right halves; how many of the newest are kept
    Half h[kKeep] = {};             // kept: h[0] newest
    for (;;) {
        const double mx = mid<HALF>(c.x0, c.x3), my = mid<HALF>(c.y0, c.y3);
        const double ax = c.x1 - mx, ay = c.y1 - my, bx = c.x2 - mx, by = c.y2 - my;
        const double d1 = fma(ax, ax, ay * ay);
        const double d2 = fma(bx, bx, by * by);
        if (size == 1 || (d1 <= flat_sq && d2 <= flat_sq)) {
            leaf(c.x3, c.y3);
            if (pending == 0) return;
            // the next node is the newest pending right half: [a, a + lowbit(a))
            --pending;
            a += size;
            size = a & -a;
            if (kept > 0) {
                c = Cubic{c.x3, c.y3, h[0].x1, h[0].y1, h[0].x2, h[0].y2, h[0].x3, h[0].y3};
#pragma unroll
                for (int k = 0; k + 1 < kKeep; ++k) h[k] = h[k + 1];
                --kept;
                continue;
            }
            c = root;
            int32_t lo = 0;
            for (int32_t s = kSlots; s > size; s >>= 1) {
                Cubic l, r;
                split<HALF>(c, l, r);
                if (a >= lo + (s >> 1)) {
                    c = r;
                    lo += s >> 1;
                } else {
                    c = l;
                }
            }
            continue;
        }
        Cubic l, r;
        split<HALF>(c, l, r);
#pragma unroll
        for (int k = kKeep - 1; k > 0; --k) h[k] = h[k - 1];
        h[0] = Half{r.x1, r.y1, r.x2, r.y2, r.x3, r.y3};
        kept = kept < kKeep ? kept + 1 : kKeep;
        ++pending;
        c = l;
        size >>= 1;
    }
}

// flatten_leaves_t with the (a + b) / 2 midpoint wherever the root allows it
// (the same values as f64::midpoint there; a root with a coordinate beyond
// f64::MAX/2, or NaN, takes the general form).
template <typename F>
__host__ __device__ __forceinline__ void flatten_leaves(const Cubic &root, double flat_sq, F &&leaf) {
    if (in_half_range(root))
        flatten_leaves_t<true>(root, flat_sq, leaf);
    else
        flatten_leaves_t<false>(root, flat_sq, leaf);
}

struct XpathArgs {
    const double *xy;       // user space, 2 per point
    const uint8_t *flags;   // 1 per point
    const int64_t *pt_off;  // n_paths + 1
    const double *matrix;   // 6 per path, or null: identity
    int64_t n_paths, n_pts;
    double flat_sq;
    double2 *t;          // transformed points
    uint8_t *role;       // kRole* per point
    long long *close_to;    // subpath start the closing line after this point goes to, or -1
    int32_t *cnt;           // per point: 1 if it emits edges; then per segment: edges it emits
    const long long *sidx;  // exclusive scan of the point flags (n_pts + 1): segment index, segment count
    long long *seg;         // segment -> point
    const long long *eoff;  // exclusive scan of the segment edge counts (n_pts + 1)
    double4 *edges;
    int64_t *edge_off;      // n_paths + 1
    double2 *stage;         // per segment, the first stage_n leaf points of its curve (see leaf_slot)
    int32_t stage_n;
    uint8_t *bkt;           // per segment: work class (kClasses)
    int32_t *cls_cnt;       // per (class, block of 256 segments): segments
    const long long *cls_off;  // exclusive scan of cls_cnt, class-major
    long long *perm;        // segments ordered by work class: the count pass's thread -> segment
    int32_t *err;           // set when a path's point range is out of order or out of bounds
};

// Roles of one 64-point chunk from its masks: lines, curve starts, points a
// closing line follows, subpath starts; carry = the subpath start in force
// when the chunk begins.
__device__ __forceinline__ void put_chunk(const XpathArgs &A, int64_t base, int64_t q1, uint64_t m_line,
                                          uint64_t m_curve, uint64_t m_close, uint64_t m_start, int64_t carry) {
    const int lane = threadIdx.x & 63;
    const int64_t i = base + lane;
    if (i >= q1) return;
    const uint8_t r = (m_line >> lane) & 1 ? kRoleLine : ((m_curve >> lane) & 1 ? kRoleCurve : kRoleNone);
    const bool cl = (m_close >> lane) & 1;
    const uint64_t upto = m_start & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const int64_t target = upto ? base + 63 - __clzll(static_cast<long long>(upto)) : carry;
    A.role[i] = r;
    A.close_to[i] = cl ? target : -1;
    A.cnt[i] = r != kRoleNone || cl;
}

// The walk's states, and the transition of one point as a byte table: byte x
// is the state after the point when the walk reaches it in state x.  Applying
// g and then f is one byte permute (byte x of the result is byte g[x] of f).
using rrdev::dpp_keep;
using rrdev::kDppRowBcast15;
using rrdev::kDppRowBcast31;
using rrdev::kDppRowShr1;
using rrdev::kDppRowShr2;
using rrdev::kDppRowShr4;
using rrdev::kDppRowShr8;
using rrdev::kDppWaveShr1;

enum : uint32_t { kOut = 0, kIn = 1, kSkip2 = 2, kSkip1 = 3 };
constexpr uint32_t kWalkIdentity = 0x03020100u;

__device__ __forceinline__ uint32_t walk_then(uint32_t f, uint32_t g) { return __builtin_amdgcn_perm(f, f, g); }

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ uint32_t walk_from(uint32_t t) {
    return static_cast<uint32_t>(
        dpp_keep<CTRL, ROW_MASK>(static_cast<int32_t>(t), static_cast<int32_t>(kWalkIdentity)));
}

// The XPath::new walk of one path (xpath.rs:174-258), one wave per path.  The
// lanes transform a chunk of 64 points (xpath.rs:421-426) and run the walk's
// state machine over the chunk's flags in parallel: each lane forms its
// point's transition table, a prefix scan of the tables (later after earlier)
// gives every lane the state the walk reaches its point in, and the roles
// follow from that state and the point's flags as 64-bit masks.  A curve
// never checks the LAST flag, so a subpath ending in a curve runs on into the
// next point; a curve without its two further points ends the subpath at the
// point before it (which may lie in the previous chunk, so a chunk's roles
// are stored one chunk late) and is looked at again outside a subpath.
__global__ void __launch_bounds__(256) k_xpath_walk(XpathArgs A) {
    const int lane = threadIdx.x & 63;
    const int64_t p = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (p >= A.n_paths) return;
    const int64_t q0 = A.pt_off[p], q1 = A.pt_off[p + 1];
    if (q0 < 0 || q1 < q0 || q1 > A.n_pts) {
        if (lane == 0) atomicOr(A.err, 1);
        return;
    }
    double m[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 0.0};
    if (A.matrix)
        for (int k = 0; k < 6; ++k) m[k] = A.matrix[6 * p + k];
    uint32_t state = kOut;  // entering the chunk
    int64_t s = -1;         // start of the subpath being walked
    uint64_t p_line = 0, p_curve = 0, p_close = 0, p_start = 0;
    int64_t p_base = -1, p_carry = -1;
    for (int64_t base = q0; base < q1; base += 64) {
        const int64_t i = base + lane;
        const bool live = i < q1;
        uint32_t f = 0;
        if (live) {
            f = A.flags[i];
            const double xi = A.xy[2 * i], yi = A.xy[2 * i + 1];
            A.t[i] = make_double2(fma(xi, m[0], fma(yi, m[2], m[4])), fma(xi, m[1], fma(yi, m[3], m[5])));
        }
        const int nk = static_cast<int>(min<int64_t>(64, q1 - base));
        const int64_t carry = s;
        const bool first = f & kFirst, curve = f & kCurve, last = f & kLast;
        const bool near_end = i + 2 >= q1;  // a curve here lacks its two further points
        const uint32_t n_out = first ? kIn : kOut;
        const uint32_t n_in = curve ? (near_end ? n_out : kSkip2) : (last ? kOut : kIn);
        uint32_t t = live ? (n_out | (n_in << 8) | (kSkip1 << 16) | (kIn << 24)) : kWalkIdentity;
        t = walk_then(t, walk_from<kDppRowShr1>(t));
        t = walk_then(t, walk_from<kDppRowShr2>(t));
        t = walk_then(t, walk_from<kDppRowShr4>(t));
        t = walk_then(t, walk_from<kDppRowShr8>(t));
        t = walk_then(t, walk_from<kDppRowBcast15, 0xa>(t));
        t = walk_then(t, walk_from<kDppRowBcast31, 0xc>(t));
        const uint32_t x = (walk_from<kDppWaveShr1>(t) >> (8 * state)) & 0xffu;  // state at this point
        const bool in = live && x == kIn;
        const bool cut = in && curve && near_end;  // the subpath ends at the point before
        const uint64_t m_start = __ballot((live && x == kOut && first) || (cut && first));
        const uint64_t m_curve = __ballot(in && curve && !near_end);
        const uint64_t m_line = __ballot(in && !curve);
        const uint64_t m_cut = __ballot(cut);
        uint64_t m_close = __ballot(in && !curve && last) | (m_cut >> 1);
        if (m_cut & 1) p_close |= 1ull << 63;
        state = (static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int32_t>(t), nk - 1)) >> (8 * state)) & 0xffu;
        if (base + 64 >= q1 && state != kOut) m_close |= 1ull << (nk - 1);  // the path ends inside a subpath
        if (m_start) s = base + 63 - __clzll(static_cast<long long>(m_start));
        if (p_base >= 0) put_chunk(A, p_base, q1, p_line, p_curve, p_close, p_start, p_carry);
        p_line = m_line;
        p_curve = m_curve;
        p_close = m_close;
        p_start = m_start;
        p_base = base;
        p_carry = carry;
    }
    if (p_base >= 0) put_chunk(A, p_base, q1, p_line, p_curve, p_close, p_start, p_carry);
}

__global__ void __launch_bounds__(256) k_xpath_compact(XpathArgs A) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < A.n_pts && A.cnt[i]) A.seg[A.sidx[i]] = i;
}

__device__ __forceinline__ Cubic curve_at(const XpathArgs &A, int64_t i) {
    const double2 p0 = A.t[i - 1], p1 = A.t[i], p2 = A.t[i + 1], p3 = A.t[i + 2];
    return Cubic{p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y};
}

__device__ __forceinline__ bool closes(const XpathArgs &A, int64_t i, double2 *cur, double2 *start) {
    const long long s = A.close_to[i];
    if (s < 0) return false;
    *cur = A.t[i];
    *start = A.t[s];
    return fabs(cur->x - start->x) > 1e-10 || fabs(cur->y - start->y) > 1e-10;
}

// Staging slot of leaf k of segment sg (a segment's leaves are contiguous, so
// the emit pass reads an edge's two points from one place).
__device__ __forceinline__ int64_t leaf_slot(const XpathArgs &A, int64_t sg, int32_t k) {
    return sg * A.stage_n + k;
}

// Work classes for the count pass: a curve's flatten work grows with how far
// its root's control points lie from the chord midpoint, relative to the
// flatness (about a quarter power of the ratio in leaves); class 2 + log2 of
// that ratio, 1 for a curve that is flat at once, 0 for a line or closing
// line.  Waves then take segments of one class, so that their lanes finish
// together.  The order only affects which thread counts which segment.
constexpr int kClasses = 64;

__device__ __forceinline__ int32_t work_class(const XpathArgs &A, int64_t i) {
    if (A.role[i] != kRoleCurve) return 0;
    const Cubic c = curve_at(A, i);
    const double mx = midpoint(c.x0, c.x3), my = midpoint(c.y0, c.y3);
    const double ax = c.x1 - mx, ay = c.y1 - my, bx = c.x2 - mx, by = c.y2 - my;
    const double d = fmax(fma(ax, ax, ay * ay), fma(bx, bx, by * by));
    if (isnan(d)) return kClasses - 1;
    if (!(d > A.flat_sq)) return 1;
    const double q = d / A.flat_sq;
    if (!(q < 1e300)) return kClasses - 1;
    return min(2 + ilogb(q), kClasses - 1);
}

__global__ void __launch_bounds__(256) k_xpath_classify(XpathArgs A, int64_t n_seg, int64_t n_blk) {
    __shared__ int32_t hist[kClasses];
    if (threadIdx.x < kClasses) hist[threadIdx.x] = 0;
    __syncthreads();
    const int64_t sg = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (sg < n_seg) {
        const int32_t k = work_class(A, A.seg[sg]);
        A.bkt[sg] = static_cast<uint8_t>(k);
        atomicAdd(&hist[k], 1);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) A.cls_cnt[threadIdx.x * n_blk + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_xpath_order(XpathArgs A, int64_t n_seg, int64_t n_blk) {
    __shared__ int32_t rank[kClasses];
    if (threadIdx.x < kClasses) rank[threadIdx.x] = 0;
    __syncthreads();
    const int64_t sg = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (sg < n_seg) {
        const int32_t k = A.bkt[sg];
        const int32_t r = atomicAdd(&rank[k], 1);
        A.perm[A.cls_off[k * n_blk + blockIdx.x] + r] = sg;
    }
}

__global__ void __launch_bounds__(256) k_xpath_count(XpathArgs A) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= A.n_pts) return;
    if (t >= A.sidx[A.n_pts]) {
        A.cnt[t] = 0;  // past the last segment: adds nothing to the offsets
        return;
    }
    const int64_t sg = A.perm[t];
    const int64_t i = A.seg[sg];
    int32_t n = 0;
    const uint8_t r = A.role[i];
    if (r == kRoleLine) {
        n = 1;
    } else if (r == kRoleCurve) {
        // count the leaves, keeping the first stage_n for the emit pass
        flatten_leaves(curve_at(A, i), A.flat_sq, [&](double x, double y) {
            if (n < A.stage_n) A.stage[leaf_slot(A, sg, n)] = make_double2(x, y);
            ++n;
        });
    }
    double2 cur, start;
    if (closes(A, i, &cur, &start)) ++n;
    A.cnt[sg] = n;
}

// Emit: a wave takes 64 consecutive segments, whose edges form one contiguous
// run of the output, and writes that run with coalesced stores: output edge q
// finds its segment by a binary search over the 65 segment offsets held in
// LDS, then takes its two points from the transformed points or the staged
// leaves.  A curve with more leaves than were staged is flattened again by
// its own lane afterwards.
constexpr uint32_t kSegCurve = 1, kSegClose = 2, kSegRedo = 4;

__device__ __forceinline__ double4 seg_edge(const XpathArgs &A, int64_t sg, int64_t i, uint32_t info, int64_t k,
                                            int64_t leaves) {
    if (info & kSegCurve) {
        if (k < leaves) {
            const double2 b = A.stage[leaf_slot(A, sg, static_cast<int32_t>(k))];
            const double2 a = k == 0 ? A.t[i - 1] : A.stage[leaf_slot(A, sg, static_cast<int32_t>(k - 1))];
            return make_double4(a.x, a.y, b.x, b.y);
        }
    } else if (A.role[i] == kRoleLine && k == 0) {
        const double2 a = A.t[i - 1], b = A.t[i];
        return make_double4(a.x, a.y, b.x, b.y);
    }
    // the closing line after point i
    const double2 cur = A.t[i], start = A.t[A.close_to[i]];
    return make_double4(cur.x, cur.y, start.x, start.y);
}

__global__ void __launch_bounds__(256) k_xpath_emit(XpathArgs A) {
    __shared__ long long off[4][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t n_seg = A.sidx[A.n_pts];
    const int64_t sg0 = (static_cast<int64_t>(blockIdx.x) * 4 + w) * 64, sg = sg0 + lane;
    const bool valid = sg < n_seg;
    // this lane's segment: point, kind, edge count (eoff is flat past n_seg)
    int64_t i = 0, leaves = 0;
    uint32_t info = 0;
    const long long o0 = A.eoff[min(sg, A.n_pts)];
    off[w][lane] = o0;
    if (lane == 63) off[w][64] = A.eoff[min(sg0 + 64, A.n_pts)];
    if (valid) {
        i = A.seg[sg];
        double2 cur, start;
        const bool cl = closes(A, i, &cur, &start);
        leaves = A.eoff[sg + 1] - o0 - (cl ? 1 : 0);
        info = (cl ? kSegClose : 0u);
        if (A.role[i] == kRoleCurve) info |= kSegCurve | (leaves > A.stage_n ? kSegRedo : 0u);
    }
    __shared__ uint32_t s_info[4][64];
    __shared__ long long s_pt[4][64], s_leaves[4][64];
    s_info[w][lane] = info;
    s_pt[w][lane] = i;
    s_leaves[w][lane] = leaves;
    __syncthreads();
    const long long base = off[w][0], end = off[w][64];
    for (long long q = base + lane; q < end; q += 64) {
        int j = 0;  // last segment j with off[j] <= q
#pragma unroll
        for (int step = 32; step > 0; step >>= 1)
            if (off[w][j + step] <= q) j += step;
        const uint32_t inf = s_info[w][j];
        if (inf & kSegRedo) continue;
        A.edges[q] = seg_edge(A, sg0 + j, s_pt[w][j], inf, q - off[w][j], s_leaves[w][j]);
    }
    if (info & kSegRedo) {
        long long o = o0;
        const Cubic c = curve_at(A, i);
        double px = c.x0, py = c.y0;
        flatten_leaves(c, A.flat_sq, [&](double x, double y) {
            A.edges[o++] = make_double4(px, py, x, y);
            px = x;
            py = y;
        });
        double2 cur, start;
        if (closes(A, i, &cur, &start)) A.edges[o] = make_double4(cur.x, cur.y, start.x, start.y);
    }
}

__global__ void __launch_bounds__(256) k_xpath_offsets(XpathArgs A) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p <= A.n_paths) A.edge_off[p] = A.eoff[A.sidx[A.pt_off[p]]];
}

#define RR_TRY(ctx, expr)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (expr);                                                                             \
        if (e_ != hipSuccess) return rrint::fail(ctx, RR_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

constexpr long long kStageLeaves = 64;         // leaves kept per curve by the count pass
constexpr long long kStageBytes = 4ll << 30;   // at most this much staging

}  // namespace

extern "C" int32_t rr_hip_xpath_edges(rr_hip_ctx *c, const rr_xpath_batch *b, int64_t *n_edges) {
    if (!c || !b || !n_edges) return RR_E_INVARIANT;
    if (rrint::page_open(c)) return rrint::fail(c, RR_E_INVARIANT, "xpath_edges inside an open page");
    if (b->n_paths < 0 || b->n_pts < 0 || !b->pt_off || !b->edge_off || (b->n_pts > 0 && (!b->xy || !b->flags)) ||
        b->edge_cap < 0 || (b->edge_cap > 0 && !b->edges) || !(b->flatness >= 0.0))
        return rrint::fail(c, RR_E_INVARIANT, "xpath_edges: bad batch");
    if (b->n_pts >= 0x7fffffffll * 256ll || b->n_paths >= 0x7fffffffll * 256ll)
        return rrint::fail(c, RR_E_INVARIANT, "xpath_edges: batch too large");
    RR_TRY(c, hipSetDevice(rrint::device(c)));
    *n_edges = 0;
    const int64_t n_pts = b->n_pts;
    int32_t rc = rrint::ensure(c, &c->xp_t, &c->xp_t_cap, static_cast<size_t>(std::max<int64_t>(n_pts, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_role, &c->xp_role_cap, static_cast<size_t>(std::max<int64_t>(n_pts, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_close, &c->xp_close_cap, static_cast<size_t>(std::max<int64_t>(n_pts, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_cnt, &c->xp_cnt_cap, static_cast<size_t>(std::max<int64_t>(n_pts, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_eoff, &c->xp_eoff_cap, static_cast<size_t>(n_pts + 1));
    if (!rc) rc = rrint::ensure(c, &c->xp_sidx, &c->xp_sidx_cap, static_cast<size_t>(n_pts + 1));
    if (!rc) rc = rrint::ensure(c, &c->xp_seg, &c->xp_seg_cap, static_cast<size_t>(std::max<int64_t>(n_pts, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_err, &c->xp_err_cap, 1);
    if (rc) return rc;
    XpathArgs A;
    A.xy = b->xy;
    A.flags = b->flags;
    A.pt_off = b->pt_off;
    A.matrix = b->matrix;
    A.n_paths = b->n_paths;
    A.n_pts = n_pts;
    A.flat_sq = b->flatness * b->flatness;
    A.t = c->xp_t;
    A.role = c->xp_role;
    A.close_to = c->xp_close;
    A.cnt = c->xp_cnt;
    A.eoff = c->xp_eoff;
    A.sidx = c->xp_sidx;
    A.seg = c->xp_seg;
    A.edges = reinterpret_cast<double4 *>(b->edges);
    A.edge_off = b->edge_off;
    A.err = c->xp_err;
    hipStream_t st = rrint::stream(c);
    const bool prof = c->profiling;
    for (int k = 0; prof && k < 6; ++k)
        if (!c->xp_ev[k]) RR_TRY(c, hipEventCreate(&c->xp_ev[k]));
    auto mark = [&](int k) -> hipError_t { return prof ? hipEventRecord(c->xp_ev[k], st) : hipSuccess; };
    c->xp_ms[0] = c->xp_ms[1] = c->xp_ms[2] = 0.0;
    // the point ranges must tile [0, n_pts): checked before any point is read
    // by the per-point kernels
    int64_t ends[2] = {0, 0};
    int32_t err = 0;
    RR_TRY(c, hipMemsetAsync(c->xp_err, 0, sizeof(int32_t), st));
    RR_TRY(c, hipMemcpyAsync(&ends[0], b->pt_off, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RR_TRY(c, hipMemcpyAsync(&ends[1], b->pt_off + b->n_paths, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    RR_TRY(c, mark(0));
    if (b->n_paths > 0) {
        hipLaunchKernelGGL(k_xpath_walk, dim3(static_cast<unsigned>((b->n_paths + 3) / 4)), dim3(256), 0, st, A);
        RR_TRY(c, hipGetLastError());
    }
    RR_TRY(c, mark(1));
    RR_TRY(c, hipMemcpyAsync(&err, c->xp_err, sizeof(err), hipMemcpyDeviceToHost, st));
    RR_TRY(c, hipStreamSynchronize(st));
    if (ends[0] != 0 || ends[1] != n_pts || err)
        return rrint::fail(c, RR_E_INVARIANT, "xpath_edges: pt_off must rise from 0 to n_pts");
    rc = rrint::scan_counts(c, c->xp_cnt, c->xp_sidx, n_pts);
    if (rc) return rc;
    long long n_seg = 0;
    RR_TRY(c, hipMemcpyAsync(&n_seg, c->xp_sidx + n_pts, sizeof(n_seg), hipMemcpyDeviceToHost, st));
    if (n_pts > 0) {
        hipLaunchKernelGGL(k_xpath_compact, dim3(blocks_for(n_pts)), dim3(256), 0, st, A);
        RR_TRY(c, hipGetLastError());
    }
    RR_TRY(c, hipStreamSynchronize(st));
    // leaf staging: up to kStageLeaves leaves per segment within kStageBytes
    const long long seg_pad = (n_seg + 63) / 64 * 64;
    A.stage_n = seg_pad > 0 ? static_cast<int32_t>(std::min<long long>(
                                  kStageLeaves, kStageBytes / (seg_pad * static_cast<long long>(sizeof(double2)))))
                            : 0;
    if (A.stage_n > 0) {
        rc = rrint::ensure(c, &c->xp_stage, &c->xp_stage_cap, static_cast<size_t>(seg_pad * A.stage_n));
        if (rc) return rc;
    }
    A.stage = c->xp_stage;
    // segments ordered by work class for the count pass
    const long long n_blk = (n_seg + 255) / 256;
    rc = rrint::ensure(c, &c->xp_bkt, &c->xp_bkt_cap, static_cast<size_t>(std::max<long long>(n_seg, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_perm, &c->xp_perm_cap, static_cast<size_t>(std::max<long long>(n_seg, 1)));
    if (!rc)
        rc = rrint::ensure(c, &c->xp_ccnt, &c->xp_ccnt_cap, static_cast<size_t>(std::max<long long>(kClasses * n_blk, 1)));
    if (!rc) rc = rrint::ensure(c, &c->xp_coff, &c->xp_coff_cap, static_cast<size_t>(kClasses * n_blk + 1));
    if (rc) return rc;
    A.bkt = c->xp_bkt;
    A.perm = c->xp_perm;
    A.cls_cnt = c->xp_ccnt;
    A.cls_off = c->xp_coff;
    if (n_seg > 0) {
        hipLaunchKernelGGL(k_xpath_classify, dim3(static_cast<unsigned>(n_blk)), dim3(256), 0, st, A,
                           static_cast<int64_t>(n_seg), static_cast<int64_t>(n_blk));
        RR_TRY(c, hipGetLastError());
        rc = rrint::scan_counts(c, c->xp_ccnt, c->xp_coff, kClasses * n_blk);
        if (rc) return rc;
        hipLaunchKernelGGL(k_xpath_order, dim3(static_cast<unsigned>(n_blk)), dim3(256), 0, st, A,
                           static_cast<int64_t>(n_seg), static_cast<int64_t>(n_blk));
        RR_TRY(c, hipGetLastError());
    }
    RR_TRY(c, mark(2));
    if (n_pts > 0) {
        hipLaunchKernelGGL(k_xpath_count, dim3(blocks_for(n_pts)), dim3(256), 0, st, A);
        RR_TRY(c, hipGetLastError());
    }
    RR_TRY(c, mark(3));
    rc = rrint::scan_counts(c, c->xp_cnt, c->xp_eoff, n_pts);
    if (rc) return rc;
    hipLaunchKernelGGL(k_xpath_offsets, dim3(blocks_for(b->n_paths + 1)), dim3(256), 0, st, A);
    RR_TRY(c, hipGetLastError());
    long long total = 0;
    RR_TRY(c, hipMemcpyAsync(&total, c->xp_eoff + n_pts, sizeof(total), hipMemcpyDeviceToHost, st));
    RR_TRY(c, hipStreamSynchronize(st));
    *n_edges = total;
    if (total > b->edge_cap)
        return rrint::fail(c, RR_E_CAPACITY, "xpath_edges: edge buffer holds " + std::to_string(b->edge_cap) +
                                                 " edges, the paths give " + std::to_string(total));
    RR_TRY(c, mark(4));
    if (n_seg > 0) {
        hipLaunchKernelGGL(k_xpath_emit, dim3(static_cast<unsigned>((n_seg + 255) / 256)), dim3(256), 0, st, A);
        RR_TRY(c, hipGetLastError());
    }
    RR_TRY(c, mark(5));
    RR_TRY(c, hipStreamSynchronize(st));
    if (prof)
        for (int k = 0; k < 3; ++k) {
            float ms = 0.f;
            RR_TRY(c, hipEventElapsedTime(&ms, c->xp_ev[2 * k], c->xp_ev[2 * k + 1]));
            c->xp_ms[k] = ms;
        }
    return RR_OK;
}

extern "C" int32_t rr_hip_last_xpath_stats(rr_hip_ctx *c, double ms[3]) {
    if (!c || !ms) return RR_E_INVARIANT;
    for (int k = 0; k < 3; ++k) ms[k] = c->xp_ms[k];
    return RR_OK;
}
