// clusters.hip - cluster-pair segments, bounds, filter, angle sort, line-fit
// moments (src/main.cpp:571-801).
//
// After the key sort, equal keys form segments (the reference's
// reduce_by_segment, src/main.cpp:571-585). Each segment's ClusterBounds is
// reduced by one wave and filtered (ValidBlobFilter, src/main.cpp:604-609);
// segments too small or too long to
// pass are rejected from their length alone. Valid segments are numbered in
// key order (their cluster index) and laid out contiguously (the extents of
// src/main.cpp:700-717).
// Every record of a valid segment is placed at its final position with its
// calc_theta slope (src/main.cpp:621-675). Then one work-group per cluster
// sorts its points by (slope, plane, y, x) - the reference's angle sort
// (src/main.cpp:737-750) made stable - in
// LDS (global scratch for clusters larger than LDS), computes each point's
// line-fit moments from the grey image (src/main.cpp:768-775) and the
// cluster's inclusive prefix sums in exact int64 (src/main.cpp:793-801).
#include "clusters.hpp"
#include "segsort.hpp"

namespace atq {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

// Segment heads: keys differ above `shift` (the detector's keys carry the
// plane in their 2 low bits).
struct HeadFlag {
    const uint64_t *keys;
    uint32_t shift;
    ATQ_DEV uint32_t operator()(uint32_t i) const {
        return i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift);
    }
};
// Segment of every record, first record of every segment, and the end
// sentinel seg_start[S] = N written by the last record.
struct HeadOut {
    uint32_t *seg_id, *seg_start;
    const uint32_t *n_dev;
    uint32_t cap;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        seg_id[i] = pre + v - 1;
        if (v) seg_start[pre] = i;
        if (i + 1 == min(*n_dev, cap)) seg_start[pre + v] = i + 1;
    }
};

struct ValidFlag {
    const uint32_t *valid;
    ATQ_DEV uint32_t operator()(uint32_t s) const { return valid[s]; }
};
struct ValidOut {
    uint32_t *seg_cluster;
    const atq_cluster_bounds *seg_bounds;
    atq_cluster_bounds *bounds;
    ATQ_DEV void operator()(uint32_t s, uint32_t pre, uint32_t v) const {
        seg_cluster[s] = v ? pre : NONE;
        if (v) bounds[pre] = seg_bounds[s];
    }
};

struct ClusterCount {
    const atq_cluster_bounds *bounds;
    ATQ_DEV uint32_t operator()(uint32_t c) const { return bounds[c].count; }
};
struct ExtentOut {
    atq_cluster_extents *ext;
    ATQ_DEV void operator()(uint32_t c, uint32_t pre, uint32_t v) const { ext[c] = {pre, v}; }
};

}  // namespace

// One work-group per segment: reduce ClusterBounds (BoundsAcc) and apply
// ValidBlobFilter. Each thread loads 8 records per round before using them.
constexpr int SB_NT = 1024, SB_IPT = 8;  // k_seg_bounds: a segment's records per round

// FX: fixed mode (both half-pixel bits decoded).
template <bool FX>
__global__ void __launch_bounds__(SB_NT)
k_seg_bounds(const atq_boundary_point *__restrict__ pts, const uint32_t *__restrict__ seg_start,
             const atq_counts *counts, uint32_t w, uint32_t h, uint32_t *__restrict__ valid,
             atq_cluster_bounds *__restrict__ seg_bounds) {
    constexpr int NW = SB_NT / 64;
    __shared__ BoundsAcc part[NW];
    const uint32_t S = counts->n_pairs;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t max_px = 2ull * (2ull * w + 2ull * h);
    for (uint32_t s = blockIdx.x; s < S; s += gridDim.x) {
        const uint32_t b = seg_start[s], e = seg_start[s + 1], c = e - b;
        if (c < 24 || c > max_px) {
            if (threadIdx.x == 0) valid[s] = 0;
            continue;
        }
        BoundsAcc acc;
        acc.init();
        // every load of a round in flight before the first is used
        for (uint32_t i0 = b + threadIdx.x; i0 < e; i0 += SB_IPT * SB_NT) {
            atq_boundary_point q[SB_IPT];
#pragma unroll
            for (int k = 0; k < SB_IPT; k++) {
                const uint32_t i = i0 + SB_NT * k;
                q[k] = i < e ? pts[i] : atq_boundary_point{0, 0};
            }
#pragma unroll
            for (int k = 0; k < SB_IPT; k++)
                if (i0 + SB_NT * k < e) acc.add(q[k], FX);
        }
        acc.wave_reduce();
        if (lane == 0) part[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 1; q < NW; q++) acc.merge(part[q]);
            atq_cluster_bounds cb;
            cb.x_min = (uint16_t)acc.xmin;
            cb.x_max = (uint16_t)acc.xmax;
            cb.y_min = (uint16_t)acc.ymin;
            cb.y_max = (uint16_t)acc.ymax;
            cb.gx_sum = acc.gxs;
            cb.gy_sum = acc.gys;
            cb.pxgx_pygy_sum = acc.pg;
            cb.start = b;
            cb.count = c;
            const bool ok = valid_blob(cb, w, h);
            valid[s] = ok;
            if (ok) seg_bounds[s] = cb;
        }
        __syncthreads();
    }
}

// Each record of a valid segment goes to its final slot with its slope.
template <bool FX>
__global__ void __launch_bounds__(256)
k_assign(const atq_boundary_point *__restrict__ pts, const uint32_t *__restrict__ seg_id,
         const uint32_t *__restrict__ seg_start, const uint32_t *__restrict__ seg_cluster,
         const atq_cluster_bounds *__restrict__ bounds, const atq_cluster_extents *__restrict__ ext,
         const uint32_t *n_dev, uint32_t cap, atq_cluster_point *__restrict__ cpts,
         uint32_t *__restrict__ cidx) {
    const uint32_t n = min(*n_dev, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t s = seg_id[i];
        uint32_t c = seg_cluster[s];
        if (c == NONE) continue;
        uint32_t pos = ext[c].start + (i - seg_start[s]);
        atq_boundary_point p = pts[i];
        const atq_cluster_bounds &b = bounds[c];
        float th = calc_theta(p.packed_x, p.packed_y, bounds_cx(b), bounds_cy(b), FX);
        cpts[pos] = atq_cluster_point{p.packed_x, p.packed_y, th};
        cidx[pos] = c;
    }
}

// Cluster assignment in the reference's own form: lower_bound over the filtered clusters
// (src/main.cpp:621-662), for callers that hold no segment table.
__global__ void __launch_bounds__(256)
k_assign_search(const atq_boundary_point *__restrict__ pts, const uint32_t *n_dev, uint32_t cap,
                const atq_cluster_bounds *__restrict__ bounds, const atq_counts *counts,
                const atq_cluster_extents *__restrict__ ext, atq_cluster_point *__restrict__ cpts,
                uint32_t *__restrict__ cidx) {
    const uint32_t n = min(*n_dev, cap);
    const uint32_t ncl = counts->n_clusters;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = ncl;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (bounds[mid].start + bounds[mid].count <= i) lo = mid + 1;
            else hi = mid;
        }
        if (lo == ncl) continue;
        const atq_cluster_bounds &b = bounds[lo];
        if (i < b.start || i >= b.start + b.count) continue;
        atq_boundary_point p = pts[i];
        uint32_t pos = ext[lo].start + (i - b.start);
        cpts[pos] = atq_cluster_point{p.packed_x, p.packed_y,
                                      calc_theta(p.packed_x, p.packed_y, bounds_cx(b), bounds_cy(b))};
        cidx[pos] = lo;
    }
}

// Extents of the runs of equal cluster index (for callers holding only the
// per-point index array, as the reference's segmented scan does).
struct RunHead {
    const uint32_t *cidx;
    ATQ_DEV uint32_t operator()(uint32_t i) const { return i == 0 || cidx[i] != cidx[i - 1]; }
};
struct RunOut {
    atq_cluster_extents *ext;
    const uint32_t *n_dev;
    uint32_t cap;
    const uint32_t *cidx;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        if (!v) return;
        const uint32_t n = min(*n_dev, cap);
        uint32_t e = i + 1;
        while (e < n && cidx[e] == cidx[i]) e++;
        ext[pre] = atq_cluster_extents{i, e - i};
    }
};

// Angle sort keys: slope << 32 | position in the cluster, so equal slopes
// keep the input order (the reference's sort, made stable).
// The angle sort's key of point j of the cluster at `start`: its slope's
// order-preserving image over its position. When cidx is set it also records
// the point's cluster index (every point's key is generated exactly once,
// by the chunk sort, and a cluster's points keep their positions' range).
struct AngleKey {
    static constexpr bool position_low = true;  // low word = index in the cluster
    const atq_cluster_point *in;
    uint32_t *cidx;
    ATQ_DEV uint64_t operator()(uint32_t c, uint32_t start, uint32_t j) const {
        if (cidx) cidx[start + j] = c;
        return ((uint64_t)float_key(in[start + j].slope) << 32) | j;
    }
};

// Prefix records staged per round of FIT_STAGE positions: the owners put
// their moments in LDS, then consecutive lanes write consecutive 16-byte
// quarters of the 64-byte records (a thread's own four records lie 256 bytes
// from its neighbours', which leaves every store instruction scattered). A
// thread's four staged records (24 int64) take FIT_STAGE_THREAD = 25 slots:
// at a stride of 24 the owners' stores hit the same few LDS banks (48 words
// apart), at 25 (50 words) they spread over the banks.
constexpr uint32_t FIT_STAGE = 1024;
constexpr uint32_t FIT_STAGE_THREAD = 6 * SEG_IPT + 1;
constexpr uint32_t FIT_STAGE_WORDS = FIT_STAGE / SEG_IPT * FIT_STAGE_THREAD;
ATQ_DEV uint32_t fit_stage_ix(uint32_t pos) {
    return FIT_STAGE_THREAD * (pos / SEG_IPT) + 6 * (pos % SEG_IPT);
}

// Loads the points of chunk positions e0 + i < m (src[i]: their positions in
// the cluster) and their line-fit terms: weight W and doubled coordinates
// RX, RY (zero past m); sum gets the thread's moment sums. Two rounds of
// independent loads (points, then their grey neighbourhoods) instead of a
// dependent chain per point.
ATQ_DEV void fit_terms(const atq_cluster_point *in, uint32_t start, uint32_t e0, uint32_t m,
                       const uint32_t (&src)[SEG_IPT], const uint8_t *grey, uint32_t w,
                       uint32_t h, atq_cluster_point (&p)[SEG_IPT], uint32_t (&W)[SEG_IPT],
                       uint32_t (&RX)[SEG_IPT], uint32_t (&RY)[SEG_IPT], int64_t (&sum)[6],
                       bool fx) {
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) p[i] = in[start + (e0 + i < m ? src[i] : 0u)];
    uint32_t g4[SEG_IPT][4];
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++)
        linefit_gather(p[i].packed_x, p[i].packed_y, w, h, grey, g4[i], fx);
#pragma unroll
    for (int q = 0; q < 6; q++) sum[q] = 0;
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) {
        W[i] = RX[i] = RY[i] = 0;
        if (e0 + i >= m) continue;
        linefit_terms_from(p[i].packed_x, p[i].packed_y, w, h, g4[i], &W[i], &RX[i], &RY[i], fx);
        const atq_line_fit_point f = linefit_from_terms(W[i], RX[i], RY[i]);
        sum[0] += f.Mx;
        sum[1] += f.My;
        sum[2] += f.Mxx;
        sum[3] += f.Myy;
        sum[4] += f.Mxy;
        sum[5] += f.W;
    }
}

// The same terms from the points already in angle order (sorted[at + e0 +
// i]) and their weights (wts, k_fit_terms): one round of coalesced loads.
ATQ_DEV void fit_terms_sorted(const atq_cluster_point *sorted, const uint16_t *wts, uint32_t at,
                              uint32_t e0, uint32_t m, uint32_t (&W)[SEG_IPT],
                              uint32_t (&RX)[SEG_IPT], uint32_t (&RY)[SEG_IPT], int64_t (&sum)[6],
                              bool fx) {
    atq_cluster_point p[SEG_IPT];
    uint32_t wt[SEG_IPT];
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) {
        const bool in = e0 + i < m;
        p[i] = in ? sorted[at + e0 + i] : atq_cluster_point{};
        wt[i] = in ? wts[at + e0 + i] : 0u;
    }
#pragma unroll
    for (int q = 0; q < 6; q++) sum[q] = 0;
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) {
        W[i] = RX[i] = RY[i] = 0;
        if (e0 + i >= m) continue;
        W[i] = wt[i];
        RX[i] = x_with_dx(p[i].packed_x, fx) + 1;
        RY[i] = y_with_dy(p[i].packed_x, p[i].packed_y, fx) + 1;
        const atq_line_fit_point f = linefit_from_terms(W[i], RX[i], RY[i]);
        sum[0] += f.Mx;
        sum[1] += f.My;
        sum[2] += f.Mxx;
        sum[3] += f.Myy;
        sum[4] += f.Mxy;
        sum[5] += f.W;
    }
}

// Slot of (cluster c, sub-chunk q) in the sub-chunk sum table: c + (start +
// q * FIT_SUB) / FIT_SUB is strictly increasing over the sub-chunks of
// successive clusters, so the table needs n_clusters + n_points / FIT_SUB
// + 1 slots.
ATQ_DEV uint32_t fit_sub_slot(uint32_t c, uint32_t start, uint32_t q) {
    return c + (start + q * FIT_SUB) / FIT_SUB;
}

// Every cluster's points in angle order to `out` and their line-fit weights
// (include/line_fit_point.h:50-86: 1 + the grey gradient's magnitude) to
// wts, FIT_SUB positions per work-group item, four per thread in lockstep
// (sorted keys, then points, then grey neighbourhoods: three rounds of
// independent loads). Few registers, so many work-groups share a CU and
// hide the gathers' latency; the prefix-sum kernels then read both arrays
// coalesced. With sub_sums, a long cluster's sub-chunks (all but those of
// its last SEG_CHUNK chunk) also leave their six moment sums there: the
// carries that the chunk-parallel prefix sums add.
template <bool FX>
__global__ void __launch_bounds__(256)
k_fit_terms(const atq_cluster_point *__restrict__ in, const uint64_t *__restrict__ keys0,
            const uint64_t *__restrict__ keys1, const atq_cluster_extents *__restrict__ ext,
            const uint32_t *n_clusters_dev, uint32_t cap_cl, uint32_t max_subs,
            const uint8_t *__restrict__ grey, uint32_t w, uint32_t h,
            atq_cluster_point *__restrict__ out, uint16_t *__restrict__ wts,
            int64_t *__restrict__ sub_sums, uint32_t lds) {
    __shared__ int64_t part[6][4];
    const uint32_t ncl = min(*n_clusters_dev, cap_cl);
    const uint64_t items = (uint64_t)ncl * max_subs;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {  // sub-chunk-major
        const uint32_t c = (uint32_t)(it % ncl), sb = (uint32_t)(it / ncl);
        const uint32_t start = ext[c].start, n = ext[c].count, b = sb * FIT_SUB;
        if (b >= n) continue;
        const uint64_t *keys = seg_sorted_keys(n, keys0, keys1, lds != 0) + start;
        uint32_t j[4], src[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            j[i] = b + threadIdx.x + 256 * i;
            src[i] = j[i] < n ? (uint32_t)keys[j[i]] : 0u;
        }
        atq_cluster_point p[4];
#pragma unroll
        for (int i = 0; i < 4; i++) p[i] = in[start + src[i]];
        uint32_t g[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++) linefit_gather(p[i].packed_x, p[i].packed_y, w, h, grey, g[i], FX);
        uint32_t W[4], rx[4], ry[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            W[i] = rx[i] = ry[i] = 0;
            if (j[i] >= n) continue;
            linefit_terms_from(p[i].packed_x, p[i].packed_y, w, h, g[i], &W[i], &rx[i], &ry[i], FX);
            out[start + j[i]] = p[i];
            wts[start + j[i]] = (uint16_t)W[i];
        }
        if (sub_sums && (b / SEG_CHUNK + 1) * SEG_CHUNK < n) {  // uniform in the work-group
            int64_t sum[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const atq_line_fit_point f = linefit_from_terms(W[i], rx[i], ry[i]);
                sum[0] += f.Mx;
                sum[1] += f.My;
                sum[2] += f.Mxx;
                sum[3] += f.Myy;
                sum[4] += f.Mxy;
                sum[5] += f.W;
            }
            const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int64_t x = wave_inclusive_sum(sum[q]);
                if (lane == 63) part[q][wave] = x;
            }
            __syncthreads();
            if (threadIdx.x < 6)
                sub_sums[6 * (size_t)fit_sub_slot(c, start, sb) + threadIdx.x] =
                    part[threadIdx.x][0] + part[threadIdx.x][1] + part[threadIdx.x][2] +
                    part[threadIdx.x][3];
            __syncthreads();
        }
    }
}

// One chunk of up to SEG_CHUNK points of a cluster in angle order, chunk
// positions [base, base + m) (src[i]: the point at chunk position e0 + i,
// e0 = threadIdx.x * SEG_IPT, for e0 + i < m): writes the points to `out`
// (if non-null) and their line-fit moments (include/line_fit_point.h:50-86),
// prefix-summed from `carry` over the chunk, to `lf` (if non-null). Called by
// all SEG_NT threads; tot gets the chunk's sums; lds holds the scan's
// scratch, stage FIT_STAGE_WORDS int64 for the store exchange. Each point
// keeps only its weight and doubled coordinates in registers; the moments
// are rebuilt from them when written.
// PRE: the points are already in `out` in angle order with their weights
// in wts (k_fit_terms), and src is unused.
template <bool PRE>
ATQ_DEV void fit_chunk(const atq_cluster_point *in, atq_cluster_point *out, uint32_t start,
                       uint32_t base, uint32_t m, const uint32_t (&src)[SEG_IPT],
                       const uint8_t *grey, uint32_t w, uint32_t h, atq_line_fit_point *lf,
                       const uint16_t *wts, const int64_t (&carry)[6], int64_t (&tot)[6],
                       int64_t *lds, int64_t *stage, bool fx) {
    const uint32_t t = threadIdx.x, e0 = t * SEG_IPT;
    uint32_t W[SEG_IPT], RX[SEG_IPT], RY[SEG_IPT];
    int64_t sum[6];
    if (!lf) {
        if (out)
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++)
                if (e0 + i < m) out[start + base + e0 + i] = in[start + src[i]];
        return;
    }
    if constexpr (PRE) {
        fit_terms_sorted(out, wts, start + base, e0, m, W, RX, RY, sum, fx);
    } else {
        atq_cluster_point p[SEG_IPT];
        fit_terms(in, start, e0, m, src, grey, w, h, p, W, RX, RY, sum, fx);
        if (out) {
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++)
                if (e0 + i < m) out[start + base + e0 + i] = p[i];
        }
    }
    int64_t acc[6];
    seg_block_scan6(sum, acc, tot, lds);
#pragma unroll
    for (int q = 0; q < 6; q++) acc[q] += carry[q] - sum[q];
    int64_t *lfw = reinterpret_cast<int64_t *>(lf + start + base);
    for (uint32_t r = 0; r * FIT_STAGE < m; r++) {
        if (e0 / FIT_STAGE == r) {
            int64_t *d = stage + fit_stage_ix(e0 - r * FIT_STAGE);
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++) {
                const atq_line_fit_point f = linefit_from_terms(W[i], RX[i], RY[i]);
                acc[0] += f.Mx;
                acc[1] += f.My;
                acc[2] += f.Mxx;
                acc[3] += f.Myy;
                acc[4] += f.Mxy;
                acc[5] += f.W;
#pragma unroll
                for (int q = 0; q < 6; q++) d[6 * i + q] = acc[q];
            }
        }
        __syncthreads();
        const uint32_t nr = min(FIT_STAGE, m - r * FIT_STAGE);
        int64_t *dst = lfw + 8 * (size_t)(r * FIT_STAGE);
#pragma unroll
        for (uint32_t k = 0; k < 4 * FIT_STAGE / SEG_NT; k++) {
            const uint32_t piece = t + SEG_NT * k, rc = piece >> 2, qt = piece & 3;
            if (rc >= nr) continue;
            longlong2 v = {0, 0};
            if (qt < 3) {
                const uint32_t sx = fit_stage_ix(rc) + 2 * qt;
                v = longlong2{stage[sx], stage[sx + 1]};
            }
            *reinterpret_cast<longlong2 *>(dst + 8 * rc + 2 * qt) = v;
        }
        __syncthreads();
    }
}

// Clusters of at most SEG_CHUNK points leave the sort with their sorted keys
// in `keys` (where seg_sorted_keys puts a cluster of that length); the fit
// runs in its own kernel, so that the sort keeps few registers and several
// work-groups share a CU.
struct AngleStore {
    uint64_t *keys;
    ATQ_DEV void operator()(uint32_t, uint32_t start, uint32_t n, const uint64_t (&v)[SEG_IPT],
                            uint64_t *) const {
        const uint32_t e0 = threadIdx.x * SEG_IPT;
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++)
            if (e0 + i < n) keys[start + e0 + i] = v[i];
    }
};

// A work-group-uniform value kept in scalar registers.
ATQ_DEV int64_t uniform64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Chunk-parallel form over every cluster (its sorted keys where
// seg_sorted_keys puts them): every chunk writes its points and its moments
// prefix-summed from the sums of the cluster's earlier chunks (the
// sub-chunk sums k_fit_terms leaves for the clusters longer than SEG_CHUNK).
// SEQ: every cluster (n_clusters_dev) in input order, one work-group per
// cluster walking its chunks and carrying the sums itself. PRE: the points
// in angle order and their weights come from k_fit_terms (out, wts).
template <bool SEQ, bool FX, bool PRE>
__global__ void __launch_bounds__(SEG_NT)
k_fit_chunks(const atq_cluster_point *__restrict__ in, atq_cluster_point *__restrict__ out,
             const uint64_t *__restrict__ keys0, const uint64_t *__restrict__ keys1,
             const atq_cluster_extents *__restrict__ ext, const uint32_t *n_clusters_dev,
             uint32_t cap_cl, SegLong lng, uint32_t max_chunks, const uint8_t *__restrict__ grey,
             uint32_t w, uint32_t h, atq_line_fit_point *__restrict__ lf,
             const int64_t *__restrict__ slot_sums, const uint16_t *__restrict__ wts,
             uint32_t lds_merge) {
    __shared__ int64_t lds[6 * SEG_NT / 64 + 6];
    __shared__ int64_t stage[FIT_STAGE_WORDS];
    const uint32_t ncl = min(*n_clusters_dev, cap_cl);
    const uint64_t items = (uint64_t)ncl * (SEQ ? 1 : max_chunks);
    const uint32_t e0 = threadIdx.x * SEG_IPT;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {  // chunk-major
        const uint32_t c = SEQ ? (uint32_t)it : (uint32_t)(it % ncl);
        const uint32_t k = SEQ ? 0u : (uint32_t)(it / ncl);
        const uint32_t start = ext[c].start, n = ext[c].count, b = k * SEG_CHUNK;
        if (b >= n) continue;
        const uint64_t *keys =
            SEQ ? nullptr : seg_sorted_keys(n, keys0, keys1, lds_merge != 0) + start;
        int64_t carry[6] = {0, 0, 0, 0, 0, 0}, tot[6];
        if (!SEQ && lf && k) {
            // the sums of the 4k sub-chunks before this chunk: one round of
            // loads spread over the threads (thread t < 1020 keeps field
            // t % 6), then per field through LDS (stage is free here)
            constexpr uint32_t SPAN = SEG_NT / 6 * 6;
            const uint32_t t = threadIdx.x, nv = 6 * k * (SEG_CHUNK / FIT_SUB);
            int64_t part = 0;
            if (t < SPAN)
                for (uint32_t x = t; x < nv; x += SPAN)
                    part += slot_sums[6 * (size_t)fit_sub_slot(c, start, x / 6) + x % 6];
            if (t < SPAN) stage[t] = part;
            __syncthreads();
            if (t < 6) {
                int64_t a = 0;
                for (uint32_t q = t; q < min(nv, SPAN); q += 6) a += stage[q];
                stage[SPAN + t] = a;
            }
            __syncthreads();
#pragma unroll
            for (int f = 0; f < 6; f++) carry[f] = stage[SPAN + f];
            __syncthreads();
        }
        for (uint32_t bb = b; bb < (SEQ ? n : b + 1); bb += SEG_CHUNK) {
            uint32_t src[SEG_IPT];
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++) {
                const uint32_t j = bb + e0 + i;
                src[i] = !PRE && j < n ? (keys ? (uint32_t)keys[j] : j) : 0u;
            }
            fit_chunk<PRE>(in, out, start, bb, min(SEG_CHUNK, n - bb), src, grey, w, h, lf, wts,
                           carry, tot, lds, stage, FX);
            if (SEQ)
#pragma unroll
                for (int q = 0; q < 6; q++) carry[q] += uniform64(tot[q]);
        }
    }
}

int launch_assign_search(hipStream_t s, const atq_boundary_point *pts, const uint32_t *n_dev,
                         uint32_t cap, const atq_cluster_bounds *bounds, const atq_counts *counts,
                         const atq_cluster_extents *ext, atq_cluster_point *cpts, uint32_t *cidx) {
    k_assign_search<<<2048, 256, 0, s>>>(pts, n_dev, cap, bounds, counts, ext, cpts, cidx);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_run_extents(hipStream_t s, const uint32_t *cidx, const uint32_t *n_dev, uint32_t cap,
                       atq_cluster_extents *ext, uint32_t *n_runs_dev, ScanCtx &sc) {
    return device_scan(s, n_dev, cap, RunHead{cidx}, RunOut{ext, n_dev, cap, cidx}, sc,
                       n_runs_dev);
}

int launch_segments(hipStream_t s, const uint64_t *keys, uint32_t key_shift, const uint32_t *n_dev,
                    atq_counts *counts, uint32_t cap, uint32_t *seg_id, uint32_t *seg_start,
                    ScanCtx &sc) {
    return device_scan(s, n_dev, cap, HeadFlag{keys, key_shift},
                       HeadOut{seg_id, seg_start, n_dev, cap}, sc, &counts->n_pairs);
}

int launch_bounds_filter(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_start,
                         atq_counts *counts, uint32_t cap, uint32_t w, uint32_t h, uint32_t *valid,
                         atq_cluster_bounds *seg_bounds, uint32_t *seg_cluster,
                         atq_cluster_bounds *bounds, ScanCtx &sc, bool fx) {
    if (fx) k_seg_bounds<true><<<1024, SB_NT, 0, s>>>(pts, seg_start, counts, w, h, valid, seg_bounds);
    else k_seg_bounds<false><<<1024, SB_NT, 0, s>>>(pts, seg_start, counts, w, h, valid, seg_bounds);
    return device_scan(s, &counts->n_pairs, cap, ValidFlag{valid},
                       ValidOut{seg_cluster, seg_bounds, bounds}, sc, &counts->n_clusters);
}

int launch_extents(hipStream_t s, const atq_cluster_bounds *bounds, atq_counts *counts, uint32_t cap,
                   atq_cluster_extents *ext, ScanCtx &sc) {
    return device_scan(s, &counts->n_clusters, cap, ClusterCount{bounds}, ExtentOut{ext},
                       sc, &counts->n_cluster_points);
}

int launch_assign(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_id,
                  const uint32_t *seg_start, const uint32_t *seg_cluster,
                  const atq_cluster_bounds *bounds, const atq_cluster_extents *ext,
                  const uint32_t *n_dev, uint32_t cap, atq_cluster_point *cpts, uint32_t *cidx,
                  bool fx) {
    if (fx)
        k_assign<true><<<2048, 256, 0, s>>>(pts, seg_id, seg_start, seg_cluster, bounds, ext,
                                            n_dev, cap, cpts, cidx);
    else
        k_assign<false><<<2048, 256, 0, s>>>(pts, seg_id, seg_start, seg_cluster, bounds, ext,
                                             n_dev, cap, cpts, cidx);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

namespace {
template <bool FX>
int angle_fit(hipStream_t s, const atq_cluster_point *in, atq_cluster_point *out,
              const atq_cluster_extents *ext, const atq_counts *counts, uint32_t cap_cl,
              uint32_t max_cluster, const uint8_t *grey, uint32_t w, uint32_t h,
              atq_line_fit_point *lf, uint64_t *keys0, uint64_t *keys1, int64_t *slot_sums,
              SegLong lng, uint32_t *cidx, uint16_t *wts, bool throughput) {
    // the sort (every cluster's keys where seg_sorted_keys puts them), then
    // the points in angle order with their weights, then the fit of every
    // chunk of every cluster
    const bool lds_merge = throughput || seg_lds_merge(max_cluster);
    int r = seg_sort(s, AngleKey{in, cidx}, AngleStore{keys1}, ext, &counts->n_clusters, cap_cl,
                     max_cluster, keys0, keys1, lng, lds_merge);
    if (r) return r;
    const uint32_t max_chunks = seg_max_chunks(max_cluster);
    const uint32_t lds = lds_merge ? 1u : 0u;
    if (lf && wts) {
        const uint32_t max_subs = (max_cluster + FIT_SUB - 1) / FIT_SUB;
        k_fit_terms<FX><<<8192, 256, 0, s>>>(in, keys0, keys1, ext, &counts->n_clusters, cap_cl,
                                             max_subs ? max_subs : 1, grey, w, h, out, wts,
                                             max_cluster > SEG_CHUNK ? slot_sums : nullptr, lds);
    }
    // one work-group per chunk slot at most a few rounds deep: the dispatcher
    // hands the chunks that exist to whichever CU is free
    if (lf)
        k_fit_chunks<false, FX, true><<<2048, SEG_NT, 0, s>>>(in, out, keys0, keys1, ext,
                                                             &counts->n_clusters, cap_cl, lng,
                                                             max_chunks, grey, w, h, lf, slot_sums,
                                                             wts, lds);
    else
        k_fit_chunks<false, FX, false><<<2048, SEG_NT, 0, s>>>(in, out, keys0, keys1, ext,
                                                              &counts->n_clusters, cap_cl, lng,
                                                              max_chunks, grey, w, h, lf,
                                                              slot_sums, nullptr, lds);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}
}  // namespace

int launch_angle_fit(hipStream_t s, const atq_cluster_point *in, atq_cluster_point *out,
                     const atq_cluster_extents *ext, const atq_counts *counts, uint32_t cap_cl,
                     uint32_t max_cluster, const uint8_t *grey, uint32_t w, uint32_t h,
                     atq_line_fit_point *lf, uint64_t *keys0, uint64_t *keys1, int64_t *slot_sums,
                     SegLong lng, bool fx, uint32_t *cidx, uint16_t *wts, bool throughput) {
    if (lf && !wts) return ATQ_ERR_BAD_ARG;
    return fx ? angle_fit<true>(s, in, out, ext, counts, cap_cl, max_cluster, grey, w, h, lf,
                                keys0, keys1, slot_sums, lng, cidx, wts, throughput)
              : angle_fit<false>(s, in, out, ext, counts, cap_cl, max_cluster, grey, w, h, lf,
                                 keys0, keys1, slot_sums, lng, cidx, wts, throughput);
}

int launch_cluster_fit(hipStream_t s, const atq_cluster_point *in, const atq_cluster_extents *ext,
                       const uint32_t *n_clusters_dev, uint32_t cap_cl, const uint8_t *grey,
                       uint32_t w, uint32_t h, atq_line_fit_point *lf) {
    k_fit_chunks<true, false, false><<<1024, SEG_NT, 0, s>>>(in, nullptr, nullptr, nullptr, ext,
                                                      n_clusters_dev, cap_cl, SegLong{}, 1, grey,
                                                      w, h, lf, nullptr, nullptr, 0u);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
