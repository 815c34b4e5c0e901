// fitquad.hip - corner candidates, quad selection and corner geometry
// (src/fit_quad.cpp, src/main.cpp:832-1162).
//
// fit_lines: the reference's 32-item work-groups with their two-cluster
// bookkeeping (src/fit_quad.cpp:153-375) compute, for every point x >= 1,
// exactly a per-cluster cyclic computation: window errors
// e(p) = n_w * eig_err(moment[p-k, p+k]) with k = min(20, n/12), the 7-tap
// filter f(p) = sum_t F[t] e(p+t-3) accumulated in order, and a strict local
// maximum test f(p) > f(p-1), f(p) > f(p+1), all indices modulo the cluster
// size n; global index 0 is never a candidate (its work-group slot is a
// halo; in fixed mode it is). Here one thread per point evaluates that
// definition directly from the prefix-summed moments.
//
// Corners (error != 0) are compacted, grouped per cluster and sorted by
// (error desc, index asc) - src/main.cpp:832-907 with a stable sort. do_indexing runs one 256-item work-group per peak
// segment like the reference (src/fit_quad.cpp:392-535), pairing segment e
// with cluster extents[e] as the reference does (fixed mode: with its own
// cluster's extents). The quad checks of
// process_fitted_quads run one thread per fitted quad.
#include "clusters.hpp"
#include "segsort.hpp"

namespace atq {

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr float MAX_LINE_FIT_MSE = 10.0f;
// (float)cos(10 deg) and cos(10 deg) as glibc computes 10*M_PI/180
constexpr float COS10_F = 0x1.f838b8p-1f;
constexpr double COS10_D = 0x1.f838b8c811c17p-1;

__constant__ float FILTER[7] = {0x1.6c0504p-7f, 0x1.152aaap-3f, 0x1.368b3p-1f, 0x1p+0f,
                                0x1.368b3p-1f,  0x1.152aaap-3f, 0x1.6c0504p-7f};

ATQ_DEV atq_line_fit_point lf_read(const atq_line_fit_point *p, int64_t i, int64_t bound,
                                   uint32_t *oob) {
    if (i < 0 || i >= bound) {
        if (oob) atomicAdd(oob, 1u);
        return atq_line_fit_point{0, 0, 0, 0, 0, 0, 0, 0};
    }
    return p[i];
}

ATQ_DEV void lf_sub(atq_line_fit_point &a, const atq_line_fit_point &b) {
    a.Mx -= b.Mx; a.My -= b.My; a.Mxx -= b.Mxx; a.Myy -= b.Myy; a.Mxy -= b.Mxy; a.W -= b.W;
}
ATQ_DEV void lf_add(atq_line_fit_point &a, const atq_line_fit_point &b) {
    a.Mx += b.Mx; a.My += b.My; a.Mxx += b.Mxx; a.Myy += b.Myy; a.Mxy += b.Mxy; a.W += b.W;
}

// src/fit_quad.cpp:15-35
ATQ_DEV atq_line_fit_point get_moment(const atq_line_fit_point *p, uint32_t n, uint32_t i0,
                                      uint32_t i1, uint32_t *total, int64_t bound, uint32_t *oob) {
    atq_line_fit_point r{0, 0, 0, 0, 0, 0, 0, 0};
    *total = 0;
    if (i0 < i1) {
        *total = i1 - i0 + 1;
        r = lf_read(p, i1, bound, oob);
        if (i0 > 0) lf_sub(r, lf_read(p, (int64_t)i0 - 1, bound, oob));
    } else if (i1 < i0) {
        *total = n - i0 + i1 + 1;
        r = lf_read(p, (int64_t)n - 1, bound, oob);
        lf_sub(r, lf_read(p, (int64_t)i0 - 1, bound, oob));
        lf_add(r, lf_read(p, i1, bound, oob));
    }
    return r;
}

// src/fit_quad.cpp:37-105 (int64 products wrap; no FMA contraction)
ATQ_DEV void fit_line(const atq_line_fit_point &m, uint32_t num, float *line, float *err,
                      float *mse) {
    int64_t Csxx = wsub(wmul(m.Mxx, m.W), wmul(m.Mx, m.Mx));
    int64_t Csxy = wsub(wmul(m.Mxy, m.W), wmul(m.Mx, m.My));
    int64_t Csyy = wsub(wmul(m.Myy, m.W), wmul(m.My, m.My));
    int64_t corr = wmul(wmul(m.W, m.W), 4);
    float sdist = hypot_f((float)wsub(Csxx, Csyy), (float)wmul(2, Csxy));
    float eig_err = ((float)wadd(Csxx, Csyy) - sdist) / (float)wmul(corr, 2);
    if (line) {
        float nx1 = (float)wsub(Csxx, Csyy) - sdist;
        float ny1 = (float)wmul(2, Csxy);
        float nx2 = ny1;
        float ny2 = (float)wsub(Csyy, Csxx) - sdist;
        float M1 = nx1 * nx1 + ny1 * ny1;
        float M2 = nx2 * nx2 + ny2 * ny2;
        float nx, ny;
        if (M1 > M2) { nx = nx1; ny = ny1; } else { nx = nx2; ny = ny2; }
        line[0] = (float)m.Mx / (float)wmul(m.W, 2);
        line[1] = (float)m.My / (float)wmul(m.W, 2);
        float length = hypot_f(nx, ny);
        if ((double)fabsf(length) < 1e-12) {
            line[2] = 0.0f;
            line[3] = 0.0f;
        } else {
            line[2] = nx / length;
            line[3] = ny / length;
        }
    }
    if (err) *err = (float)num * eig_err;
    if (mse) *mse = eig_err;
}

ATQ_DEV float window_error(const atq_line_fit_point *p, uint32_t n, uint32_t k, uint32_t pos) {
    uint32_t i0 = (pos + n - k) % n, i1 = (pos + k) % n, total;
    atq_line_fit_point m = get_moment(p, n, i0, i1, &total, n, nullptr);
    float e;
    fit_line(m, total, nullptr, &e, nullptr);
    return e;
}

// Binomials for the combinatorial number system decode (n <= 10, k <= 4);
// for these the reference's uint8_t table (include/combinadics.h) is exact.
ATQ_DEV uint32_t binom(uint32_t n, uint32_t k) {
    uint32_t r = 1;
    if (k > n) return 0;
    for (uint32_t i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

// include/combinadics.h:43-59: digits c0 > c1 > c2 > c3 with
// num = C(c0,4) + C(c1,3) + C(c2,2) + C(c3,1).
ATQ_DEV void combo_decode(uint32_t num, uint32_t out[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t kk = 4 - i, c = kk - 1;
        while (binom(c + 1, kk) <= num) c++;
        out[i] = c;
        num -= binom(c, kk);
    }
}

// The quads' compaction scan's flag: fitted quad e's corners to tmp[e]
// (read back by QuadOut in the same thread) and whether it passes.
struct QuadFlag {
    const atq_fitted_quad *fitted;
    atq_quad_corners *tmp;
    int min_tag_width;
    ATQ_DEV uint32_t operator()(uint32_t e) const;
};
struct QuadOut {
    const atq_quad_corners *tmp;
    atq_quad_corners *out;
    ATQ_DEV void operator()(uint32_t e, uint32_t pre, uint32_t v) const {
        if (v) out[pre] = tmp[e];
    }
};

}  // namespace

__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ void k_zero_u32(uint32_t *p, const uint32_t *n_dev, uint32_t cap, uint32_t set_idx,
                           uint32_t set_val) {
    uint32_t n = n_dev ? min(*n_dev, cap) : cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = i == set_idx ? set_val : 0u;
}

// One thread per cluster point: the cyclic window / filter / maximum.
// Error of the window of 2k+1 points centred on every cluster point
// (k = min(20, n/12); 0 when k < 2).
__global__ void __launch_bounds__(256)
k_window_err(const atq_line_fit_point *__restrict__ lf, const uint32_t *__restrict__ cidx,
             const atq_cluster_extents *__restrict__ ext, const uint32_t *n_dev, uint32_t cap,
             float *__restrict__ errs) {
    const uint32_t N = n_dev ? min(*n_dev, cap) : cap;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x) {
        const atq_cluster_extents e = ext[cidx[x]];
        const uint32_t n = e.count, k = min(20u, n / 12u);
        errs[x] = k >= 2 ? window_error(lf + e.start, n, k, x - e.start) : 0.0f;
    }
}

// Per point: the filtered errors at positions j-1, j, j+1 (cyclic in the
// cluster, 7-tap filter over 9 window errors) and a corner where the middle
// one is a strict maximum. Window errors come from `errs` (k_window_err) or,
// when it is null, are computed here. Points below global index `first`
// (1 as in the reference, 0 in fixed mode) are never corners.
__global__ void __launch_bounds__(256)
k_fit_lines(const atq_line_fit_point *__restrict__ lf, const uint32_t *__restrict__ cidx,
            const atq_cluster_extents *__restrict__ ext, const uint32_t *n_dev, uint32_t cap,
            const float *__restrict__ errs, atq_corner *__restrict__ sparse,
            uint32_t *__restrict__ ccount, uint32_t first) {
    const uint32_t N = n_dev ? min(*n_dev, cap) : cap;
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x) {
        const uint32_t c = cidx[x];
        const atq_cluster_extents e = ext[c];
        const uint32_t s = e.start, n = e.count, j = x - s;
        const uint32_t k = min(20u, n / 12u);
        atq_corner out{0, 0, 0.0f};
        if (x >= first && k >= 2) {
            const atq_line_fit_point *p = lf + s;
            float er[9];
#pragma unroll
            for (int d = 0; d < 9; d++) {
                const uint32_t pos = (j + n + d - 4) % n;
                er[d] = errs ? errs[s + pos] : window_error(p, n, k, pos);
            }
            float f[3];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < 7; t++) acc = acc + er[q + t] * FILTER[t];
                f[q] = acc;
            }
            if (f[1] > f[0] && f[1] > f[2]) {
                out = atq_corner{j, c, f[1]};
                if (f[1] != 0.0f && ccount) atomicAdd(&ccount[c], 1u);
            }
        }
        sparse[x] = out;
    }
}

// Corner grouping without atomics: the sparse corners (one slot per cluster
// point, in cluster order) are compacted by a scan on error != 0, the runs
// of equal cluster index become the peak segments, and each segment is
// sorted by (error desc, index asc) - the reference's order made stable,
// since the index is unique within a cluster.
namespace {
struct CornerFlag {
    const atq_corner *sparse;
    ATQ_DEV uint32_t operator()(uint32_t i) const { return sparse[i].error != 0.0f; }
};
struct CornerOut {
    const atq_corner *sparse;
    atq_corner *out;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        if (v) out[pre] = sparse[i];
    }
};
struct SegHead {
    const atq_corner *c;
    ATQ_DEV uint32_t operator()(uint32_t i) const {
        return i == 0 || c[i].cluster_index != c[i - 1].cluster_index;
    }
};
struct SegStart {
    atq_peak_extents *peaks;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        if (v) peaks[pre].start = i;
    }
};
}  // namespace

__global__ void k_peak_close(const atq_counts *counts, uint32_t cap, atq_peak_extents *peaks) {
    const uint32_t E = min(counts->n_segments, cap), nc = counts->n_corners;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < E; e += gridDim.x * blockDim.x) {
        const uint32_t end = e + 1 < E ? peaks[e + 1].start : nc;
        peaks[e].count = end - peaks[e].start;
    }
}

// Peak order inside a segment: error descending, then line-fit index
// ascending (the reference's stable sort by error, include/fit_quad.h).
struct CornerKey {
    const atq_corner *c;
    ATQ_DEV uint64_t operator()(uint32_t, uint32_t start, uint32_t j) const {
        const atq_corner cr = c[start + j];
        return ((uint64_t)~float_key(cr.error) << 32) | cr.line_fit_point_index;
    }
};

// Segments of at most SEG_CHUNK corners are written back in key order
// inside the sort, from the registers.
struct CornerFinish {
    atq_corner *out;
    ATQ_DEV void operator()(uint32_t, uint32_t start, uint32_t n, const uint64_t (&v)[SEG_IPT],
                            uint64_t *) const {
        const uint32_t c = out[start].cluster_index;
        __syncthreads();
        const uint32_t e0 = threadIdx.x * SEG_IPT;
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++)
            if (e0 + i < n)
                out[start + e0 + i] =
                    atq_corner{(uint32_t)v[i], c, float_from_key(~(uint32_t)(v[i] >> 32))};
    }
};

// Writes the corners of the listed long segments back in key order. The
// cluster index is read from the segment's first corner, which every writer
// of that slot rewrites with the same cluster index.
__global__ void __launch_bounds__(256)
k_corner_emit(const atq_peak_extents *__restrict__ peaks, SegLong lng, uint32_t max_chunks,
              const uint64_t *__restrict__ keys0, const uint64_t *__restrict__ keys1,
              atq_corner *out, uint32_t lds) {
    const uint32_t nl = min(*lng.n, lng.cap);
    const uint64_t items = (uint64_t)nl * max_chunks;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {  // chunk-major
        const uint32_t e = lng.list[it % nl], ch = (uint32_t)(it / nl);
        const uint32_t cs = peaks[e].start, m = peaks[e].count;
        if (ch * SEG_CHUNK >= m) continue;
        const uint32_t c = out[cs].cluster_index;
        const uint64_t *k = seg_sorted_keys(m, keys0, keys1, lds != 0) + cs;
        const uint32_t end = min(m, (ch + 1) * SEG_CHUNK);
        constexpr int PER = SEG_CHUNK / 256;
        uint64_t kk[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {  // all loads in flight before the stores
            const uint32_t r = ch * SEG_CHUNK + threadIdx.x + 256 * q;
            kk[q] = r < end ? k[r] : 0;
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const uint32_t r = ch * SEG_CHUNK + threadIdx.x + 256 * q;
            if (r < end)
                out[cs + r] = atq_corner{(uint32_t)kk[q], c, float_from_key(~(uint32_t)(kk[q] >> 32))};
        }
    }
}

// do_indexing: one 256-item work-group per peak segment.
__global__ void __launch_bounds__(256)
k_do_indexing(const atq_peak_extents *__restrict__ peaks, const uint32_t *n_seg_dev, uint32_t cap,
              const atq_corner *__restrict__ corners, const atq_line_fit_point *__restrict__ lf,
              const uint32_t *n_points_dev, uint32_t n_points_host,
              const atq_cluster_extents *__restrict__ ext, atq_fitted_quad *__restrict__ fitted,
              uint32_t *oob, uint32_t fx) {
    __shared__ float red_f[4];
    __shared__ uint32_t red_u[4];
    const uint32_t E = n_seg_dev ? min(*n_seg_dev, cap) : cap;
    const uint32_t NP = n_points_dev ? *n_points_dev : n_points_host;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (uint32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const atq_peak_extents pk = peaks[e];
        // paired by position as the reference does, or (fixed mode) by cluster
        const atq_cluster_extents ce = ext[fx ? corners[pk.start].cluster_index : e];
        const int64_t bound = (int64_t)NP - (int64_t)ce.start;
        const atq_line_fit_point *p = lf + ce.start;
        const uint32_t ncomb = binom(min(pk.count, 10u), 4);
        const bool valid = t < ncomb;
        bool valid_fit = valid;
        float total_err = 0.0f;
        atq_fitted_quad q;
        uint32_t li[4] = {0, 0, 0, 0};
        if (valid) {
            uint32_t ind[4];
            combo_decode(t, ind);
#pragma unroll
            for (int k = 0; k < 4; k++) li[k] = corners[pk.start + ind[k]].line_fit_point_index;
            auto sw = [&](int a, int b) {
                if (!(li[a] < li[b])) { uint32_t tmp = li[a]; li[a] = li[b]; li[b] = tmp; }
            };
            sw(0, 1); sw(2, 3); sw(0, 2); sw(1, 3); sw(1, 2);  // src/fit_quad.cpp:377-389
            float p01[4], p12[4];
#pragma unroll
            for (int sgi = 0; sgi < 4; sgi++) {
                uint32_t total;
                atq_line_fit_point m = get_moment(p, ce.count, li[sgi], li[(sgi + 1) & 3], &total, bound, oob);
                float err, mse;
                fit_line(m, total, sgi == 0 ? p01 : (sgi == 1 ? p12 : nullptr), &err, &mse);
                valid_fit = valid_fit && (mse <= MAX_LINE_FIT_MSE);
                q.moments[sgi] = m;
                q.num_in_moments[sgi] = (uint16_t)total;
                total_err = total_err + err;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) q.indices[k] = (uint16_t)li[k];
            float dot = p01[2] * p12[2] + p01[3] * p12[3];
            valid_fit = valid_fit && (fabsf(dot) <= COS10_F);
        }
        const float cur = valid_fit ? total_err : __builtin_inff();
        // work-group minimum
        float mn = cur;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, o, 64));
        if (lane == 0) red_f[wave] = mn;
        __syncthreads();
        const float lowest = fminf(fminf(red_f[0], red_f[1]), fminf(red_f[2], red_f[3]));
        // smallest combination number within 0.1 of the minimum
        uint32_t cand = ((double)(cur - lowest) < 0.1) ? t : NONE;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) cand = min(cand, (uint32_t)__shfl_xor(cand, o, 64));
        if (lane == 0) red_u[wave] = cand;
        __syncthreads();
        const uint32_t chosen = min(min(red_u[0], red_u[1]), min(red_u[2], red_u[3]));
        if (lowest == __builtin_inff()) {
            if (t == 0) {
                atq_fitted_quad z;
                for (int k = 0; k < 4; k++) {
                    z.moments[k] = atq_line_fit_point{0, 0, 0, 0, 0, 0, 0, 0};
                    z.num_in_moments[k] = 0;
                    z.indices[k] = 0;
                }
                fitted[e] = z;
            }
        } else if (t == chosen) {
            for (int k = 0; k < 4; k++) {
                q.moments[k].a = 0;
                q.moments[k].b = 0;
            }
            fitted[e] = q;
        }
        __syncthreads();
    }
}

// process_fitted_quads (src/main.cpp:1035-1162) minus decoding, for fitted
// quad e: its corners to *out, returns whether it passes.
ATQ_DEV uint32_t quad_of(const atq_fitted_quad &q, uint32_t e, int min_tag_width,
                         atq_quad_corners *out) {
    uint32_t ok = 0;
    atq_quad_corners c;
    c.reversed_border = 0;
    c.pad_[0] = c.pad_[1] = c.pad_[2] = 0;
    c.blob_index = e;
    for (int i = 0; i < 4; i++) c.corners[i][0] = c.corners[i][1] = 0.0f;
    if (q.indices[0] | q.indices[1] | q.indices[2] | q.indices[3]) {
        ok = 1;
        float lines[4][4];
        for (int i = 0; i < 4; i++) {
            float err, mse;
            fit_line(q.moments[i], q.num_in_moments[i], lines[i], &err, &mse);
        }
        for (int i = 0; i < 4 && ok; i++) {
            const int i1 = (i + 1) & 3;
            double A00 = lines[i][3], A01 = -lines[i1][3];
            double A10 = -lines[i][2], A11 = lines[i1][2];
            double B0 = -lines[i][0] + lines[i1][0];
            double B1 = -lines[i][1] + lines[i1][1];
            double det = A00 * A11 - A10 * A01;
            double W00 = A11 / det, W01 = -A01 / det;
            if (fabs(det) < 0.001) { ok = 0; break; }
            double L0 = W00 * B0 + W01 * B1;
            c.corners[i][0] = (float)((double)lines[i][0] + L0 * A00);
            c.corners[i][1] = (float)((double)lines[i][1] + L0 * A10);
        }
        if (ok) {
            float area = 0.0f, len[3], pp;
            for (int i = 0; i < 3; i++) {
                int a = i, b = (i + 1) % 3;
                len[i] = hypot_f(c.corners[b][0] - c.corners[a][0], c.corners[b][1] - c.corners[a][1]);
            }
            pp = (len[0] + len[1] + len[2]) / 2;
            area += sqrtf(pp * (pp - len[0]) * (pp - len[1]) * (pp - len[2]));
            const int idxs[4] = {2, 3, 0, 2};
            for (int i = 0; i < 3; i++) {
                int a = idxs[i], b = idxs[i + 1];
                len[i] = hypot_f(c.corners[b][0] - c.corners[a][0], c.corners[b][1] - c.corners[a][1]);
            }
            pp = (len[0] + len[1] + len[2]) / 2;
            area += sqrtf(pp * (pp - len[0]) * (pp - len[1]) * (pp - len[2]));
            if ((double)area < 0.95 * min_tag_width * min_tag_width) ok = 0;
        }
        for (int i = 0; i < 4 && ok; i++) {
            int i0 = i, i1 = (i + 1) & 3, i2 = (i + 2) & 3;
            float dx1 = c.corners[i1][0] - c.corners[i0][0];
            float dy1 = c.corners[i1][1] - c.corners[i0][1];
            float dx2 = c.corners[i2][0] - c.corners[i1][0];
            float dy2 = c.corners[i2][1] - c.corners[i1][1];
            float cosd = (dx1 * dx2 + dy1 * dy2) / sqrtf((dx1 * dx1 + dy1 * dy1) * (dx2 * dx2 + dy2 * dy2));
            if ((double)fabsf(cosd) > COS10_D || dx1 * dy2 < dy1 * dx2) ok = 0;
        }
    }
    *out = c;
    return ok;
}

ATQ_DEV uint32_t QuadFlag::operator()(uint32_t e) const {
    return quad_of(fitted[e], e, min_tag_width, &tmp[e]);
}

int launch_set_u32(hipStream_t s, uint32_t *p, uint32_t v) {
    k_set_u32<<<1, 1, 0, s>>>(p, v);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_zero_u32(hipStream_t s, uint32_t *p, const uint32_t *n_dev, uint32_t cap,
                    uint32_t set_idx, uint32_t set_val) {
    k_zero_u32<<<min(1024u, atq_div_up(cap, 256) + 1), 256, 0, s>>>(p, n_dev, cap, set_idx, set_val);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_fit_lines(hipStream_t s, const atq_line_fit_point *lf, const uint32_t *cidx,
                     const atq_cluster_extents *ext, const uint32_t *n_points_dev, uint32_t cap,
                     atq_corner *sparse, uint32_t *ccount, float *errs, bool fx) {
    unsigned grid = min(4096u, atq_div_up(cap, 256) + 1);
    if (errs) k_window_err<<<grid, 256, 0, s>>>(lf, cidx, ext, n_points_dev, cap, errs);
    k_fit_lines<<<grid, 256, 0, s>>>(lf, cidx, ext, n_points_dev, cap, errs, sparse, ccount,
                                     fx ? 0u : 1u);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

// The detector's peak order. do_indexing (src/fit_quad.cpp:445-519) reads
// only a segment's min(10, n) highest-error peaks, so the detector does not
// sort whole segments: it puts those peaks first, in the reference's order
// (CornerKey ascending), moves the peaks they displace from the first
// min(10, n) places into the places they left (both in position order) and
// leaves every other peak where the compaction put it. One work-group per
// segment (which also closes the segment's extent, as k_peak_close does):
// each thread keeps the ten smallest keys of its positions (t, t + 256, ...,
// loaded four at a time), each wave draws its ten smallest from its lanes'
// lists, and wave 0 ranks the four waves' forty and plans the moves.
constexpr int CORNER_TOP = 10;
__global__ void __launch_bounds__(256)
k_corner_top(atq_peak_extents *__restrict__ peaks, const atq_counts *counts, uint32_t cap,
             atq_corner *__restrict__ c) {
    constexpr int NW = 256 / 64;
    __shared__ uint64_t wk[NW * CORNER_TOP];
    __shared__ uint32_t wp[NW * CORNER_TOP];
    __shared__ uint32_t tp[CORNER_TOP];
    __shared__ uint32_t mv_src[2 * CORNER_TOP], mv_dst[2 * CORNER_TOP];
    __shared__ uint32_t n_mv;
    const uint32_t E = min(counts->n_segments, cap), nc = counts->n_corners;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (uint32_t e = blockIdx.x; e < E; e += gridDim.x) {
        const uint32_t start = peaks[e].start;
        const uint32_t n = (e + 1 < E ? peaks[e + 1].start : nc) - start;
        if (t == 0) peaks[e].count = n;
        const uint32_t m = min(n, (uint32_t)CORNER_TOP);
        uint64_t bk[CORNER_TOP];
        uint32_t bp[CORNER_TOP];
#pragma unroll
        for (int k = 0; k < CORNER_TOP; k++) {
            bk[k] = SEG_PAD;
            bp[k] = 0;
        }
        for (uint32_t b0 = 0; b0 < n; b0 += 4 * 256) {
            atq_corner cr[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {  // the four loads in flight together
                const uint32_t i = b0 + t + 256 * q;
                cr[q] = i < n ? c[start + i] : atq_corner{};
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t i = b0 + t + 256 * q;
                uint64_t key = i < n ? ((uint64_t)~float_key(cr[q].error) << 32) |
                                           cr[q].line_fit_point_index
                                     : SEG_PAD;
                uint32_t pos = i;
#pragma unroll
                for (int k = 0; k < CORNER_TOP; k++) {  // the key sinks to its place
                    const bool lt = key < bk[k];
                    const uint64_t ck = lt ? bk[k] : key;
                    const uint32_t cp = lt ? bp[k] : pos;
                    bk[k] = lt ? key : bk[k];
                    bp[k] = lt ? pos : bp[k];
                    key = ck;
                    pos = cp;
                }
            }
        }
        // the wave's ten smallest: each round takes the smallest head
        for (int r = 0; r < CORNER_TOP; r++) {
            const uint64_t mk = wave_min_u64(bk[0]);
            const uint64_t who = __builtin_amdgcn_ballot_w64(bk[0] == mk);
            const int l = (int)__builtin_ctzll(who);
            const uint32_t p = lane_value(bp[0], l);
            if (lane == 0) {
                wk[wave * CORNER_TOP + r] = mk;
                wp[wave * CORNER_TOP + r] = p;
            }
            if ((int)lane == l) {
#pragma unroll
                for (int k = 0; k + 1 < CORNER_TOP; k++) {
                    bk[k] = bk[k + 1];
                    bp[k] = bp[k + 1];
                }
                bk[CORNER_TOP - 1] = SEG_PAD;
            }
        }
        __syncthreads();
        if (wave == 0) {
            // rank of each of the forty (keys are distinct; pads rank >= m)
            if (lane < NW * CORNER_TOP) {
                const uint64_t k0 = wk[lane];
                uint32_t rank = 0;
#pragma unroll
                for (int j = 0; j < NW * CORNER_TOP; j++) rank += wk[j] < k0 ? 1u : 0u;
                if (rank < m) tp[rank] = wp[lane];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            // moves: the ten to the front (lane k); a front place q that
            // holds none of the ten is displaced (its index among those, j)
            // into the j-th smallest of the places the ten left behind
            uint32_t mine = 0, in_top = 0;
            bool left = false;
            uint32_t left_rank = 0;
            if (lane < m) {
                mine = tp[lane];
                left = mine >= m;
#pragma unroll
                for (int k = 0; k < CORNER_TOP; k++) {
                    if ((uint32_t)k >= m) break;
                    const uint32_t o = tp[k];
                    in_top |= o == lane ? 1u : 0u;
                    left_rank += (o >= m && o < mine) ? 1u : 0u;
                }
                mv_src[lane] = mine;
                mv_dst[lane] = lane;
            }
            const uint64_t disp = __builtin_amdgcn_ballot_w64(lane < m && !in_top);
            const uint64_t below = (1ull << lane) - 1;
            if (lane < m && !in_top) mv_src[m + __popcll(disp & below)] = lane;
            if (lane < m && left) mv_dst[m + left_rank] = mine;
            if (lane == 0) n_mv = m + (uint32_t)__popcll(disp);
        }
        __syncthreads();
        // every source read before any place is overwritten
        const uint32_t nm = n_mv;
        atq_corner v{};
        if (t < nm) v = c[start + mv_src[t]];
        __syncthreads();
        if (t < nm) c[start + mv_dst[t]] = v;
        __syncthreads();
    }
}

// The corners (error != 0) of the sparse slots in point order, the form of
// the compaction scan that reads coalesced: work-group p takes slots
// [p T, (p + 1) T), T = 256 ITEMS, thread t the slots t, t + 256, ...;
// offsets from per-wave ballots in (item, wave, lane) order and a decoupled
// look-back over the work-groups (dispatched in order). *n_out = the count.
template <int ITEMS>
__global__ void __launch_bounds__(256)
k_corner_compact(const atq_corner *__restrict__ sparse, const uint32_t *n_dev, uint32_t cap,
                 atq_corner *__restrict__ compacted, uint32_t *n_out, uint64_t *flags,
                 uint64_t *next_flags, uint32_t n_flags, uint32_t *error) {
    constexpr int NW = 256 / 64;
    __shared__ uint32_t cnt[ITEMS][NW];
    __shared__ uint32_t s_prefix;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n_flags; i += gridDim.x * 256)
        next_flags[i] = 0;
    const uint32_t N = min(*n_dev, cap);
    constexpr uint32_t T = 256 * ITEMS;
    const uint32_t p = blockIdx.x, b = p * T;
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (b >= N) {
        if (p == 0 && t == 0) *n_out = 0;
        return;
    }
    atq_corner c[ITEMS];
    uint64_t bal[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {  // every load in flight together
        const uint32_t x = b + t + 256 * k;
        c[k] = x < N ? sparse[x] : atq_corner{0, 0, 0.0f};
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        bal[k] = __builtin_amdgcn_ballot_w64(c[k].error != 0.0f);
        if (lane == 0) cnt[k][wave] = (uint32_t)__popcll(bal[k]);
    }
    __syncthreads();
    uint32_t tile_sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
#pragma unroll
        for (int w = 0; w < NW; w++) tile_sum += cnt[k][w];
    if (t < 64) {
        const uint32_t excl = lb_lookback(flags, p, tile_sum, error);
        if (t == 0) s_prefix = excl;
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
    if (t == 0 && b + T >= N) *n_out = prefix + tile_sum;
    uint32_t before = prefix;
    const uint64_t below = (1ull << lane) - 1;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        uint32_t at = before;
#pragma unroll
        for (int w = 0; w < NW; w++) at += (uint32_t)w < wave ? cnt[k][w] : 0u;
        if (c[k].error != 0.0f) compacted[at + (uint32_t)__popcll(bal[k] & below)] = c[k];
#pragma unroll
        for (int w = 0; w < NW; w++) before += cnt[k][w];
    }
}

int launch_corner_segments(hipStream_t s, const atq_corner *sparse, const uint32_t *n_sparse_dev,
                           uint32_t cap, atq_counts *counts, uint32_t cap_cl, uint32_t max_seg,
                           atq_peak_extents *peaks, atq_corner *compacted, uint64_t *scratch,
                           SegLong lng, ScanCtx &sc, bool top_only) {
    int r = ATQ_OK;
    constexpr int CC_ITEMS = 8;
    const uint32_t tiles = cap / (256 * CC_ITEMS) + 1;
    if (sc.flags && tiles <= sc.flag_count) {
        uint64_t *cur = sc.flags + (size_t)(sc.calls & 1) * sc.flag_count;
        uint64_t *next = sc.flags + (size_t)((sc.calls + 1) & 1) * sc.flag_count;
        sc.calls++;
        k_corner_compact<CC_ITEMS><<<tiles, 256, 0, s>>>(sparse, n_sparse_dev, cap, compacted,
                                                         &counts->n_corners, cur, next,
                                                         sc.flag_count, sc.error);
    } else {
        r = device_scan(s, n_sparse_dev, cap, CornerFlag{sparse}, CornerOut{sparse, compacted}, sc,
                        &counts->n_corners);
    }
    if (r) return r;
    r = device_scan(s, &counts->n_corners, cap, SegHead{compacted}, SegStart{peaks}, sc,
                    &counts->n_segments);
    if (r) return r;
    if (top_only) {  // closes the extents itself
        k_corner_top<<<min(cap_cl, 2048u), 256, 0, s>>>(peaks, counts, cap_cl, compacted);
        return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
    }
    k_peak_close<<<min(64u, atq_div_up(cap_cl, 256) + 1), 256, 0, s>>>(counts, cap_cl, peaks);
    r = seg_sort(s, CornerKey{compacted}, CornerFinish{compacted}, peaks, &counts->n_segments,
                 cap_cl, max_seg, scratch, scratch + cap, lng, seg_lds_merge(max_seg));
    if (r || max_seg <= SEG_CHUNK) return r;
    k_corner_emit<<<256, 256, 0, s>>>(peaks, lng, seg_max_chunks(max_seg), scratch, scratch + cap,
                                      compacted, seg_lds_merge(max_seg) ? 1u : 0u);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_do_indexing(hipStream_t s, const atq_peak_extents *peaks, const uint32_t *n_seg_dev,
                       uint32_t cap, const atq_corner *corners, const atq_line_fit_point *lf,
                       const uint32_t *n_points_dev, uint32_t n_points_host,
                       const atq_cluster_extents *ext, atq_fitted_quad *fitted,
                       uint32_t *oob_counter, bool fx) {
    k_do_indexing<<<1024, 256, 0, s>>>(peaks, n_seg_dev, cap, corners, lf, n_points_dev,
                                       n_points_host, ext, fitted, oob_counter, fx ? 1u : 0u);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_quads(hipStream_t s, const atq_fitted_quad *fitted, const uint32_t *n_fitted_dev,
                 uint32_t cap, atq_quad_corners *tmp, atq_quad_corners *out,
                 uint32_t *n_out_dev, ScanCtx &sc, int min_tag_width) {
    return device_scan<QuadFlag, QuadOut, true>(s, n_fitted_dev, cap,
                                                QuadFlag{fitted, tmp, min_tag_width},
                                                QuadOut{tmp, out}, sc, n_out_dev);
}

}  // namespace atq
