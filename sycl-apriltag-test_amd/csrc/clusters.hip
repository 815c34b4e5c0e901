// clusters.hip - cluster-pair segments, bounds, filter, angle sort, line-fit
// moments (rows A10..A16; src/main.cpp:571-801).
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
struct HeadOut {
    uint32_t *seg_id, *seg_start;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        seg_id[i] = pre + v - 1;
        if (v) seg_start[pre] = i;
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

// seg_start[S] = N (end sentinel)
__global__ void k_seg_close(const uint32_t *n_dev, const atq_counts *counts, uint32_t cap,
                            uint32_t *seg_start) {
    uint32_t n = min(*n_dev, cap);
    seg_start[counts->n_pairs] = n;
}

// One wave per segment: reduce ClusterBounds and apply ValidBlobFilter.
__global__ void __launch_bounds__(256)
k_seg_bounds(const atq_boundary_point *__restrict__ pts, const uint32_t *__restrict__ seg_start,
             const atq_counts *counts, uint32_t w, uint32_t h, uint32_t *__restrict__ valid,
             atq_cluster_bounds *__restrict__ seg_bounds) {
    const uint32_t S = counts->n_pairs;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint64_t max_px = 2ull * (2ull * w + 2ull * h);
    for (uint32_t s = wave; s < S; s += nwaves) {
        const uint32_t b = seg_start[s], e = seg_start[s + 1], c = e - b;
        if (c < 24 || c > max_px) {
            if (lane == 0) valid[s] = 0;
            continue;
        }
        uint32_t xmin = 0xFFFF, xmax = 0, ymin = 0xFFFF, ymax = 0;
        int32_t gxs = 0, gys = 0;
        int64_t pg = 0;
        for (uint32_t i = b + lane; i < e; i += 64) {
            atq_boundary_point p = pts[i];
            uint32_t xx = x_with_dx(p.packed_x), yy = y_with_dy(p.packed_x, p.packed_y);
            int gx = grad_x(p.packed_x, p.packed_y), gy = grad_y(p.packed_x, p.packed_y);
            xmin = min(xmin, xx); xmax = max(xmax, xx);
            ymin = min(ymin, yy); ymax = max(ymax, yy);
            gxs += gx; gys += gy;
            pg += (int64_t)xx * gx + (int64_t)yy * gy;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            xmin = min(xmin, (uint32_t)__shfl_xor(xmin, o, 64));
            xmax = max(xmax, (uint32_t)__shfl_xor(xmax, o, 64));
            ymin = min(ymin, (uint32_t)__shfl_xor(ymin, o, 64));
            ymax = max(ymax, (uint32_t)__shfl_xor(ymax, o, 64));
            gxs += __shfl_xor(gxs, o, 64);
            gys += __shfl_xor(gys, o, 64);
            pg += __shfl_xor(pg, o, 64);
        }
        if (lane == 0) {
            atq_cluster_bounds cb;
            cb.x_min = (uint16_t)xmin; cb.x_max = (uint16_t)xmax;
            cb.y_min = (uint16_t)ymin; cb.y_max = (uint16_t)ymax;
            cb.gx_sum = gxs; cb.gy_sum = gys; cb.pxgx_pygy_sum = pg;
            cb.start = b; cb.count = c;
            bool ok = valid_blob(cb, w, h);
            valid[s] = ok;
            if (ok) seg_bounds[s] = cb;
        }
    }
}

// Each record of a valid segment goes to its final slot with its slope.
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
        float th = calc_theta(p.packed_x, p.packed_y, bounds_cx(b), bounds_cy(b));
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

// Angle sort: keys are slope << 32 | position in the cluster, so equal
// slopes keep the input order (the reference's sort, made stable). Every
// cluster is cut into chunks of ANGLE_CHUNK points, each sorted in LDS by one
// work-group; clusters longer than a chunk are then finished by merge passes
// in which every key finds its rank in the partner run by binary search
// (runs double each pass; a cluster needing q passes ends in buffer q % 2).
constexpr uint32_t ANGLE_CHUNK = 4096;

__global__ void __launch_bounds__(256)
k_angle_chunks(const atq_cluster_point *__restrict__ in, const atq_cluster_extents *__restrict__ ext,
               const uint32_t *n_clusters_dev, uint32_t cap_cl, uint32_t max_chunks,
               uint64_t *__restrict__ keys_out) {
    __shared__ uint64_t keys[ANGLE_CHUNK];
    const uint32_t ncl = min(*n_clusters_dev, cap_cl);
    const uint64_t items = (uint64_t)ncl * max_chunks;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t c = (uint32_t)(it / max_chunks), k = (uint32_t)(it % max_chunks);
        const uint32_t s0 = ext[c].start, n = ext[c].count;
        const uint32_t b = k * ANGLE_CHUNK;
        if (b >= n) continue;
        const uint32_t m = min(ANGLE_CHUNK, n - b);
        for (uint32_t j = threadIdx.x; j < m; j += 256)
            keys[j] = ((uint64_t)float_key(in[s0 + b + j].slope) << 32) | (b + j);
        __syncthreads();
        block_bitonic_sort<uint64_t>(keys, m, 256, false);
        for (uint32_t j = threadIdx.x; j < m; j += 256) keys_out[s0 + b + j] = keys[j];
        __syncthreads();
    }
}

ATQ_DEV uint32_t lower_bound_u64(const uint64_t *a, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One merge pass over runs of length `run` (clusters longer than `run`).
__global__ void __launch_bounds__(256)
k_angle_merge(const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout,
              const uint32_t *__restrict__ cidx, const atq_cluster_extents *__restrict__ ext,
              const uint32_t *n_points_dev, uint32_t cap, uint32_t run) {
    const uint32_t N = min(*n_points_dev, cap);
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x) {
        const atq_cluster_extents e = ext[cidx[x]];
        if (e.count <= run) continue;
        const uint32_t j = x - e.start;
        const uint32_t base = j / (2 * run) * (2 * run);
        const uint64_t *c = kin + e.start;
        const uint64_t key = c[j];
        uint32_t pos;
        if (j - base < run) {  // left run: count the right run's smaller keys
            const uint32_t rb = min(base + run, e.count), re = min(base + 2 * run, e.count);
            pos = j + lower_bound_u64(c + rb, re - rb, key);
        } else {               // right run: count the left run's smaller keys
            pos = base + (j - base - run) + lower_bound_u64(c + base, run, key);
        }
        kout[e.start + pos] = key;
    }
}

ATQ_DEV uint32_t merge_passes(uint32_t n) {
    uint32_t q = 0;
    for (uint32_t r = ANGLE_CHUNK; r < n; r <<= 1) q++;
    return q;
}

// Line-fit points + per-cluster inclusive moment prefix, one work-group per
// cluster, reading the points in angle order. `keys0/keys1` null: the input
// is already in order. `out` null: no points written. `lf` null: no moments.
// Each thread takes FIT_ITEMS consecutive points, so one round of the
// work-group covers FIT_ITEMS * 256 points with one scan of the six sums.
constexpr int FIT_ITEMS = 8;

__global__ void __launch_bounds__(256)
k_cluster_fit(const atq_cluster_point *__restrict__ in, atq_cluster_point *__restrict__ out,
              const uint64_t *__restrict__ keys0, const uint64_t *__restrict__ keys1,
              const atq_cluster_extents *__restrict__ ext, const uint32_t *n_clusters_dev,
              uint32_t cap_cl, const uint8_t *__restrict__ grey, uint32_t w, uint32_t h,
              atq_line_fit_point *__restrict__ lf) {
    __shared__ int64_t wsum[4][6];
    const uint32_t ncl = min(*n_clusters_dev, cap_cl);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t c = blockIdx.x; c < ncl; c += gridDim.x) {
        const uint32_t s0 = ext[c].start, n = ext[c].count;
        const uint64_t *keys = keys0 ? ((merge_passes(n) & 1) ? keys1 : keys0) + s0 : nullptr;
        int64_t carry[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t base = 0; base < n; base += FIT_ITEMS * 256) {
            const uint32_t j0 = base + threadIdx.x * FIT_ITEMS;
            uint32_t W[FIT_ITEMS], RX[FIT_ITEMS], RY[FIT_ITEMS];
            int64_t v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < FIT_ITEMS; k++) {
                W[k] = 0;
                RX[k] = RY[k] = 0;
                const uint32_t j = j0 + k;
                if (j < n) {
                    const uint32_t src = keys ? (uint32_t)keys[j] : j;
                    const atq_cluster_point p = in[s0 + src];
                    if (out) out[s0 + j] = p;
                    if (lf) {
                        linefit_terms(p.packed_x, p.packed_y, w, h, grey, &W[k], &RX[k], &RY[k]);
                        const atq_line_fit_point m = linefit_from_terms(W[k], RX[k], RY[k]);
                        v[0] += m.Mx;
                        v[1] += m.My;
                        v[2] += m.Mxx;
                        v[3] += m.Myy;
                        v[4] += m.Mxy;
                        v[5] += m.W;
                    }
                }
            }
            if (!lf) continue;
            // inclusive wave scan of the thread totals, then across the 4 waves
            int64_t x[6];
#pragma unroll
            for (int f = 0; f < 6; f++) {
                x[f] = v[f];
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    int64_t y = __shfl_up(x[f], o, 64);
                    if (lane >= o) x[f] += y;
                }
            }
            if (lane == 63)
#pragma unroll
                for (int f = 0; f < 6; f++) wsum[wave][f] = x[f];
            __syncthreads();
            int64_t acc[6];
#pragma unroll
            for (int f = 0; f < 6; f++) {
                int64_t before = 0, tot = 0;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int64_t t = wsum[q][f];
                    if (q < wave) before += t;
                    tot += t;
                }
                acc[f] = carry[f] + before + x[f] - v[f];
                carry[f] += tot;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < FIT_ITEMS; k++) {
                const uint32_t j = j0 + k;
                if (j < n) {
                    const atq_line_fit_point m = linefit_from_terms(W[k], RX[k], RY[k]);
                    acc[0] += m.Mx;
                    acc[1] += m.My;
                    acc[2] += m.Mxx;
                    acc[3] += m.Myy;
                    acc[4] += m.Mxy;
                    acc[5] += m.W;
                    lf[s0 + j] = atq_line_fit_point{acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], 0, 0};
                }
            }
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
                       atq_cluster_extents *ext, uint32_t *n_runs_dev, uint32_t *scan_scratch) {
    return device_scan(s, n_dev, cap, RunHead{cidx}, RunOut{ext, n_dev, cap, cidx}, scan_scratch,
                       n_runs_dev);
}

int launch_segments(hipStream_t s, const uint64_t *keys, uint32_t key_shift, const uint32_t *n_dev,
                    atq_counts *counts, uint32_t cap, uint32_t *seg_id, uint32_t *seg_start,
                    uint32_t *scan_scratch) {
    int r = device_scan(s, n_dev, cap, HeadFlag{keys, key_shift}, HeadOut{seg_id, seg_start},
                        scan_scratch, &counts->n_pairs);
    if (r) return r;
    k_seg_close<<<1, 1, 0, s>>>(n_dev, counts, cap, seg_start);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_bounds_filter(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_start,
                         atq_counts *counts, uint32_t cap, uint32_t w, uint32_t h, uint32_t *valid,
                         atq_cluster_bounds *seg_bounds, uint32_t *seg_cluster,
                         atq_cluster_bounds *bounds, uint32_t *scan_scratch) {
    k_seg_bounds<<<2048, 256, 0, s>>>(pts, seg_start, counts, w, h, valid, seg_bounds);
    return device_scan(s, &counts->n_pairs, cap, ValidFlag{valid},
                       ValidOut{seg_cluster, seg_bounds, bounds}, scan_scratch, &counts->n_clusters);
}

int launch_extents(hipStream_t s, const atq_cluster_bounds *bounds, atq_counts *counts, uint32_t cap,
                   atq_cluster_extents *ext, uint32_t *scan_scratch) {
    return device_scan(s, &counts->n_clusters, cap, ClusterCount{bounds}, ExtentOut{ext},
                       scan_scratch, &counts->n_cluster_points);
}

int launch_assign(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_id,
                  const uint32_t *seg_start, const uint32_t *seg_cluster,
                  const atq_cluster_bounds *bounds, const atq_cluster_extents *ext,
                  const uint32_t *n_dev, uint32_t cap, atq_cluster_point *cpts, uint32_t *cidx) {
    k_assign<<<2048, 256, 0, s>>>(pts, seg_id, seg_start, seg_cluster, bounds, ext, n_dev, cap,
                                  cpts, cidx);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_angle_sort(hipStream_t s, const atq_cluster_point *in, const atq_cluster_extents *ext,
                      const uint32_t *cidx, const atq_counts *counts, uint32_t cap, uint32_t cap_cl,
                      uint32_t max_cluster, uint64_t *keys0, uint64_t *keys1) {
    const uint32_t max_chunks = (max_cluster + ANGLE_CHUNK - 1) / ANGLE_CHUNK;
    k_angle_chunks<<<2048, 256, 0, s>>>(in, ext, &counts->n_clusters, cap_cl, max_chunks, keys0);
    const unsigned grid = min(4096u, atq_div_up(cap, 256) + 1);
    int pass = 0;
    for (uint32_t run = ANGLE_CHUNK; run < max_cluster; run <<= 1, pass++)
        k_angle_merge<<<grid, 256, 0, s>>>((pass & 1) ? keys1 : keys0, (pass & 1) ? keys0 : keys1,
                                           cidx, ext, &counts->n_cluster_points, cap, run);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_cluster_fit(hipStream_t s, const atq_cluster_point *in, atq_cluster_point *out,
                       const uint64_t *keys0, const uint64_t *keys1, const atq_cluster_extents *ext,
                       const uint32_t *n_clusters_dev, uint32_t cap_cl, const uint8_t *grey,
                       uint32_t w, uint32_t h, atq_line_fit_point *lf) {
    k_cluster_fit<<<2048, 256, 0, s>>>(in, out, keys0, keys1, ext, n_clusters_dev, cap_cl, grey, w,
                                       h, lf);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
