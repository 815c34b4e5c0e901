// sort.hip - LSD radix sort of (u64 key, u32 value) records (src/main.cpp:515-518).
//
// Groups the boundary records by cluster-pair key (the reference's
// sort_by_key, src/main.cpp:515-518). A pair key is min_label << 32 |
// max_label with labels below W*H; it is packed to min_label << lb |
// max_label (lb = bit length of W*H - 1), which keeps the order, so only
// 2*lb bits are sorted: 4 passes of 10-11 bit digits at 1080p, 5 at 12 MP
// (the reference's key has 64).
//
// The records are cut into G contiguous chunks of a multiple of 2048
// records, G <= RS_GROUPS (both derived from the record count on the
// device). One pass = three kernels:
//   k_rs_hist     per-chunk digit histogram (LDS atomics) -> hist[g][digit]
//   k_rs_scan     per digit: exclusive scan of its G counts in place (16
//                 waves share a digit's rows), and the digit's total
//   k_rs_scatter  each work-group first turns the digit totals into digit
//                 bases (a 2048-entry work-group scan), then in rounds of
//                 2048 records (8 per lane) each wave ranks its 512 records
//                 digit by digit with ballots (a 64-lane multisplit), the 4
//                 waves' counts are scanned in LDS, and every record is
//                 written at digit base + chunk offset + wave offset + rank.
// Records keep their input order within a digit, so the sort is stable.
#include "records.hpp"
#include "sort.hpp"

namespace atq {

namespace {
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;                       // records per lane per round
constexpr int RS_ROUND = RS_THREADS * RS_ITEMS;   // 2048
constexpr int RS_WAVE_SPAN = 64 * RS_ITEMS;       // 512 records per wave per round
}  // namespace

// Chunk size and count for n records.
ATQ_DEV void rs_chunks(uint32_t n, uint32_t *chunk, uint32_t *groups) {
    uint32_t c = max((uint32_t)RS_ROUND, (n + RS_GROUPS - 1) / RS_GROUPS);
    c = (c + RS_ROUND - 1) / RS_ROUND * RS_ROUND;
    *chunk = c;
    *groups = (n + c - 1) / c;
}

// Digit histogram of keys[b, e) into hist_row[0..2^BITS) (h: LDS of 2^BITS).
template <int BITS>
ATQ_DEV void rs_hist_body(const uint64_t *__restrict__ keys, uint32_t b, uint32_t e,
                          uint32_t shift, uint32_t *__restrict__ hist_row, uint32_t *h) {
    constexpr int B = 1 << BITS;
    for (int i = threadIdx.x; i < B; i += RS_THREADS) h[i] = 0;
    __syncthreads();
    // RS_ITEMS independent loads per lane before the LDS atomics
    for (uint32_t r = b; r < e; r += RS_ROUND) {
        uint32_t d[RS_ITEMS];
#pragma unroll
        for (int j = 0; j < RS_ITEMS; j++) {
            const uint32_t i = r + j * RS_THREADS + threadIdx.x;
            d[j] = i < e ? (uint32_t)(keys[i] >> shift) & (B - 1) : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int j = 0; j < RS_ITEMS; j++)
            if (d[j] != 0xFFFFFFFFu) atomicAdd(&h[d[j]], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < B; d += RS_THREADS) hist_row[d] = h[d];
}

template <int BITS>
__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist(const uint64_t *__restrict__ keys, const uint32_t *count, uint32_t cap, uint32_t shift,
          uint32_t *__restrict__ hist) {
    constexpr int B = 1 << BITS;
    __shared__ uint32_t h[B];
    const uint32_t n = min(*count, cap);
    uint32_t chunk, G;
    rs_chunks(n, &chunk, &G);
    if (blockIdx.x >= G) return;
    const uint32_t b = blockIdx.x * chunk, e = min(n, b + chunk);
    rs_hist_body<BITS>(keys, b, e, shift, hist + (size_t)blockIdx.x * B, h);
}

// Exclusive scan of every digit's column hist[0..G)[d] in place, and the
// digit's total to totals[d]. A work-group of 16 waves takes 64 digits (one
// per lane, so each wave reads 256 contiguous bytes of a row); wave q takes
// the q-th sixteenth of the rows: it sums them, the 16 partial sums are
// scanned in LDS, and it rescans them from its offset.
constexpr int RS_SCAN_WAVES = 16;
__global__ void __launch_bounds__(64 * RS_SCAN_WAVES)
k_rs_scan(uint32_t *__restrict__ hist, uint32_t nbuckets, const uint32_t *count, uint32_t cap,
          uint32_t *__restrict__ totals, const uint32_t *passes = nullptr, uint32_t pass = 0) {
    __shared__ uint32_t part[RS_SCAN_WAVES][64];
    if (passes && pass >= *passes) return;  // rank form: this pass is not needed
    const uint32_t n = min(*count, cap);
    uint32_t chunk, G;
    rs_chunks(n, &chunk, &G);
    const uint32_t lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const uint32_t d = blockIdx.x * 64 + lane;
    const uint32_t per = (G + RS_SCAN_WAVES - 1) / RS_SCAN_WAVES;
    const uint32_t g0 = min(G, q * per), g1 = min(G, g0 + per);
    uint32_t *col = hist + d;
    uint32_t sum = 0;
    for (uint32_t g = g0; g < g1; g += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = g + k < g1 ? col[(size_t)(g + k) * nbuckets] : 0u;
#pragma unroll
        for (int k = 0; k < 8; k++) sum += v[k];
    }
    part[q][lane] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
#pragma unroll
    for (int k = 0; k < RS_SCAN_WAVES; k++) {
        const uint32_t p = part[k][lane];
        if (k < (int)q) run += p;
        total += p;
    }
    for (uint32_t g = g0; g < g1; g += 8) {
        uint32_t v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = g + k < g1 ? col[(size_t)(g + k) * nbuckets] : 0u;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (g + k < g1) col[(size_t)(g + k) * nbuckets] = run;
            run += v[k];
        }
    }
    if (q == 0) totals[d] = total;
}

// The last rank pass's extras (on: keys hold ranks, key >> shift).
struct RsFinal {
    bool on;
    bool fx;
    uint32_t shift;
    const uint64_t *xlat;
    const float2 *centre;
    atq_cluster_point *cpts;
};

// LDS of the scatter body.
template <int BITS>
struct RsScatterLds {
    uint32_t wc[4][1 << BITS];     // per-wave digit counts, then wave offsets
    uint32_t base_off[1 << BITS];  // running output offset of each digit
    uint32_t lds4[4];
};

// Scatters records [b, e) (chunk g of the pass) to their stable positions:
// digit base (exclusive scan of `totals`) + hist_row[digit] (the digit's
// count in the chunks before this one) + rank within the chunk. fin: when
// not null (the last rank pass), keys hold ranks r (key >> rank_shift) and
// leave as (xlat[r] << 2 | plane), and each record also becomes the cluster point of
// cluster r at its position (src/main.cpp:621-675; its cluster index is
// recorded by the angle sort).
template <int BITS>
ATQ_DEV void rs_scatter_body(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                             uint64_t *__restrict__ kout, uint32_t *__restrict__ vout,
                             uint32_t b, uint32_t e, uint32_t shift,
                             const uint32_t *__restrict__ hist_row,
                             const uint32_t *__restrict__ totals, const RsFinal fin,
                             RsScatterLds<BITS> &L) {
    constexpr int B = 1 << BITS;
    constexpr int PER = B / RS_THREADS;  // digits per thread in the base scan
    auto &wc = L.wc;
    auto &base_off = L.base_off;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // the first round's records are loaded before the digit bases are
    // built, so the two rounds of loads overlap
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    auto load_round = [&](uint32_t r) {
        const uint32_t wbase = r + wave * RS_WAVE_SPAN;
#pragma unroll
        for (int j = 0; j < RS_ITEMS; j++) {
            const uint32_t i = wbase + j * 64 + lane;
            k[j] = i < e ? kin[i] : 0ull;
            v[j] = i < e ? vin[i] : 0u;
        }
    };
    load_round(b);
    {   // digit bases: exclusive scan of the totals, PER consecutive digits per thread
        uint32_t tv[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            tv[k] = totals[t * PER + k];
            sum += tv[k];
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan<uint32_t>(sum, L.lds4, &all);
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int d = t * PER + k;
            base_off[d] = run + hist_row[d];
            run += tv[k];
        }
    }
    const uint64_t lt = lanemask_lt();
    for (uint32_t r = b; r < e; r += RS_ROUND) {
        if (r != b) load_round(r);
        for (int d = t; d < B; d += RS_THREADS) {
            wc[0][d] = 0;
            wc[1][d] = 0;
            wc[2][d] = 0;
            wc[3][d] = 0;
        }
        __syncthreads();
        uint32_t dg[RS_ITEMS], rank[RS_ITEMS];
        const uint32_t wbase = r + wave * RS_WAVE_SPAN;
#pragma unroll
        for (int j = 0; j < RS_ITEMS; j++) {
            const uint32_t i = wbase + j * 64 + lane;
            const bool valid = i < e;
            const uint32_t d = (uint32_t)(k[j] >> shift) & (B - 1);
            dg[j] = valid ? d : 0xFFFFFFFFu;
            uint64_t m = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < BITS; bit++) {
                uint64_t bb = __ballot((d >> bit) & 1u);
                m &= ((d >> bit) & 1u) ? bb : ~bb;
            }
            // The lowest lane of each digit group reads and advances the
            // wave's running count of that digit and shares the old value.
            const int leader = valid ? __ffsll((unsigned long long)m) - 1 : lane;
            uint32_t before = 0;
            if (valid && lane == leader) {
                before = wc[wave][d];
                wc[wave][d] = before + (uint32_t)__popcll(m);
            }
            before = __shfl(before, leader, 64);
            rank[j] = before + (uint32_t)__popcll(m & lt);
        }
        // the last pass's per-cluster loads, in flight during the offsets below
        uint64_t xk[RS_ITEMS];
        float2 cc[RS_ITEMS];
        if (fin.on) {
#pragma unroll
            for (int j = 0; j < RS_ITEMS; j++) {
                const uint32_t r = dg[j] != 0xFFFFFFFFu ? (uint32_t)(k[j] >> fin.shift) : 0u;
                xk[j] = fin.xlat[r];
                cc[j] = fin.centre[r];
            }
        }
        __syncthreads();
        for (int d = t; d < B; d += RS_THREADS) {
            uint32_t run = base_off[d];
#pragma unroll
            for (int w = 0; w < 4; w++) {
                uint32_t c = wc[w][d];
                wc[w][d] = run;
                run += c;
            }
            base_off[d] = run;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RS_ITEMS; j++) {
            if (dg[j] != 0xFFFFFFFFu) {
                uint32_t pos = wc[wave][dg[j]] + rank[j];
                if (fin.on) {
                    kout[pos] = (xk[j] << 2) | ((v[j] >> 14) & 3u);  // the record's plane
                    const uint16_t px = (uint16_t)v[j], py = (uint16_t)(v[j] >> 16);
                    fin.cpts[pos] =
                        atq_cluster_point{px, py, calc_theta(px, py, cc[j].x, cc[j].y, fin.fx)};
                } else {
                    kout[pos] = k[j];
                }
                vout[pos] = v[j];
            }
        }
        __syncthreads();
    }
}

template <int BITS>
__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
             uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, const uint32_t *count,
             uint32_t cap, uint32_t shift, const uint32_t *__restrict__ hist,
             const uint32_t *__restrict__ totals) {
    constexpr int B = 1 << BITS;
    __shared__ RsScatterLds<BITS> L;
    const uint32_t n = min(*count, cap);
    uint32_t chunk, G;
    rs_chunks(n, &chunk, &G);
    if (blockIdx.x >= G) return;
    const uint32_t b = blockIdx.x * chunk, e = min(n, b + chunk);
    rs_scatter_body<BITS>(kin, vin, kout, vout, b, e, shift, hist + (size_t)blockIdx.x * B,
                          totals, RsFinal{}, L);
}

// Pair key min<<32|max <-> min<<lb|max.
__global__ void k_pack_keys(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                            const uint32_t *count, uint32_t cap, uint32_t lb) {
    uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k = in[i];
        out[i] = ((k >> 32) << lb) | (k & 0xFFFFFFFFull);
    }
}
__global__ void k_unpack_keys(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                              const uint32_t *count, uint32_t cap, uint32_t lb) {
    uint32_t n = min(*count, cap);
    const uint64_t lo = (1ull << lb) - 1;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t k = in[i];
        out[i] = ((k >> lb) << 32) | (k & lo);
    }
}

// ---- the detector's sort: rank keys, device-chosen pass count -------------
//
// The detector's keys are (r << 2 | plane) with r the rank of the record's
// cluster pair among the frame's valid pairs, so RK_BITS-bit digits sort
// them in counts->sort_passes passes (1 or 2; 0 = the full-key fallback
// below does the sort). Buffers X and Y ping-pong so that the last pass
// writes Y: a pass reads Y when an even number of passes, itself included,
// remain (the filter wrote its records there), else X.
ATQ_DEV bool rs_dyn_io(const RsDyn &d, uint32_t pass, const uint64_t **ki, const uint32_t **vi,
                       uint64_t **ko, uint32_t **vo, bool *last) {
    const uint32_t np = *d.passes;
    if (pass >= np) return false;
    const bool from_y = ((np - pass) & 1u) == 0;
    *ki = from_y ? d.ky : d.kx;
    *vi = from_y ? d.vy : d.vx;
    *ko = from_y ? d.kx : d.ky;
    *vo = from_y ? d.vx : d.vy;
    *last = pass + 1 == np;
    return true;
}

__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist_dyn(RsDyn d, const uint32_t *count, uint32_t cap, uint32_t pass,
              uint32_t *__restrict__ hist) {
    constexpr int B = 1 << RK_BITS;
    __shared__ uint32_t h[B];
    const uint64_t *ki;
    const uint32_t *vi;
    uint64_t *ko;
    uint32_t *vo;
    bool last;
    if (!rs_dyn_io(d, pass, &ki, &vi, &ko, &vo, &last)) return;
    const uint32_t n = min(*count, cap);
    uint32_t chunk, G;
    rs_chunks(n, &chunk, &G);
    if (blockIdx.x >= G) return;
    const uint32_t b = blockIdx.x * chunk, e = min(n, b + chunk);
    rs_hist_body<RK_BITS>(ki, b, e, pass * RK_BITS, hist + (size_t)blockIdx.x * B, h);
}

// Grid-wide barrier of a kernel whose work-groups are all resident (a grid
// of RS_FULL_GROUPS work-groups, far below what the device holds at once).
// bar counts arrivals (zero at the kernel's start); the wait is bounded:
// past LB_SPIN_LIMIT polls *error gets OVF_SCAN and the results are void.
ATQ_DEV void rs_grid_sync(uint32_t *bar, uint32_t target, uint32_t *error) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        atomicAdd(bar, 1u);
        uint32_t spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins >= LB_SPIN_LIMIT) {
                atomicOr(error, OVF_SCAN);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence();
    }
    __syncthreads();
}

// Work-group g's share [*b, *e) of n items split over G work-groups.
ATQ_DEV void rs_share(uint32_t n, uint32_t G, uint32_t g, uint32_t *b, uint32_t *e) {
    const uint32_t per = (n + G - 1) / G;
    *b = min(n, g * per);
    *e = min(n, *b + per);
}

// Exclusive scan of v over the items [b, e) of work-group g of G (v(i) in
// {0, 1, ...}); o(i, prefix, value) for each; part: G words; returns the
// total over all work-groups. Contains grid barriers: every work-group calls
// it.
template <typename V, typename O>
ATQ_DEV uint32_t rs_grid_scan(uint32_t b, uint32_t e, uint32_t G, V v, O o, uint32_t *part,
                              uint32_t *bar, uint32_t &gen, uint32_t *error, uint32_t *lds4) {
    const uint32_t g = blockIdx.x, t = threadIdx.x;
    uint32_t own = 0;
    for (uint32_t i = b + t; i < e; i += RS_THREADS) own += v(i);
    own = block_reduce_sum<uint32_t>(own, lds4);
    if (t == 0) part[g] = own;
    rs_grid_sync(bar, ++gen * G, error);
    uint32_t pre = 0, all = 0;
    for (uint32_t q = t; q < G; q += RS_THREADS) {
        const uint32_t x = part[q];
        if (q < g) pre += x;
        all += x;
    }
    pre = block_reduce_sum<uint32_t>(pre, lds4);
    all = block_reduce_sum<uint32_t>(all, lds4);
    for (uint32_t i0 = b; i0 < e; i0 += RS_THREADS) {
        const uint32_t i = i0 + t;
        const uint32_t x = i < e ? v(i) : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<uint32_t>(x, lds4, &tot);
        if (i < e) o(i, pre + ex, x);
        pre += tot;
    }
    return all;
}

// The fallback's cluster stage over the n sorted records in Y (packed keys):
// segments of equal pair (src/main.cpp:571-585), each segment's
// ClusterBounds and ValidBlobFilter (src/main.cpp:604-609), the valid ones
// numbered in key order with their extents (src/main.cpp:700-717), and each
// of their records placed as a cluster point with its slope
// (src/main.cpp:621-675).
ATQ_DEV void rs_fallback_clusters(const RsDyn &d, uint32_t n, uint32_t G, uint32_t *bar,
                                  uint32_t &gen, uint32_t *error, uint32_t *lds4,
                                  BoundsAcc *wpart) {
    const RsClusters &c = d.cl;
    const uint32_t g = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const bool fx = c.fx != 0;
    const uint64_t *K = d.ky;
    const atq_boundary_point *P = reinterpret_cast<const atq_boundary_point *>(d.vy);
    uint32_t b, e;
    // segments
    rs_share(n, G, g, &b, &e);
    auto head = [&](uint32_t i) -> uint32_t { return i == 0 || (K[i] >> 2) != (K[i - 1] >> 2); };
    const uint32_t S = rs_grid_scan(b, e, G, head, [&](uint32_t i, uint32_t pre, uint32_t v) {
        c.seg_id[i] = pre + v - 1;
        if (v) c.seg_start[pre] = i;
        if (i + 1 == n) c.seg_start[pre + v] = n;
    }, c.part, bar, gen, error, lds4);
    if (g == 0 && t == 0) c.counts->n_pairs = S;
    rs_grid_sync(bar, ++gen * G, error);
    // bounds and filter, one work-group per segment
    const uint64_t max_px = 2ull * (2ull * c.w + 2ull * c.h);
    for (uint32_t s = g; s < S; s += G) {
        const uint32_t sb = c.seg_start[s], se = c.seg_start[s + 1], cnt = se - sb;
        if (cnt < 24 || cnt > max_px) {
            if (t == 0) c.valid[s] = 0;
            continue;
        }
        BoundsAcc acc;
        acc.init();
        for (uint32_t i = sb + t; i < se; i += RS_THREADS) acc.add(P[i], fx);
        acc.wave_reduce();
        if (lane == 0) wpart[wave] = acc;
        __syncthreads();
        if (t == 0) {
            for (int q = 1; q < RS_THREADS / 64; q++) acc.merge(wpart[q]);
            atq_cluster_bounds cb;
            cb.x_min = (uint16_t)acc.xmin;
            cb.x_max = (uint16_t)acc.xmax;
            cb.y_min = (uint16_t)acc.ymin;
            cb.y_max = (uint16_t)acc.ymax;
            cb.gx_sum = acc.gxs;
            cb.gy_sum = acc.gys;
            cb.pxgx_pygy_sum = acc.pg;
            cb.start = sb;
            cb.count = cnt;
            const bool ok = valid_blob(cb, c.w, c.h);
            c.valid[s] = ok;
            if (ok) c.seg_bounds[s] = cb;
        }
        __syncthreads();
    }
    rs_grid_sync(bar, ++gen * G, error);
    // valid segments numbered in key order
    rs_share(S, G, g, &b, &e);
    const uint32_t C = rs_grid_scan(b, e, G, [&](uint32_t s) { return c.valid[s]; },
        [&](uint32_t s, uint32_t pre, uint32_t v) {
            c.seg_cluster[s] = v ? pre : 0xFFFFFFFFu;
            if (v) c.bounds[pre] = c.seg_bounds[s];
        }, c.part + G, bar, gen, error, lds4);
    if (g == 0 && t == 0) c.counts->n_clusters = C;
    rs_grid_sync(bar, ++gen * G, error);
    // extents
    rs_share(C, G, g, &b, &e);
    const uint32_t NP = rs_grid_scan(b, e, G, [&](uint32_t q) { return c.bounds[q].count; },
        [&](uint32_t q, uint32_t pre, uint32_t v) { c.ext[q] = atq_cluster_extents{pre, v}; },
        c.part + 2 * G, bar, gen, error, lds4);
    if (g == 0 && t == 0) c.counts->n_cluster_points = NP;
    rs_grid_sync(bar, ++gen * G, error);
    // cluster points
    for (uint32_t i = g * RS_THREADS + t; i < n; i += G * RS_THREADS) {
        const uint32_t s = c.seg_id[i], cl = c.seg_cluster[s];
        if (cl == 0xFFFFFFFFu) continue;
        const uint32_t pos = c.ext[cl].start + (i - c.seg_start[s]);
        const atq_boundary_point p = P[i];
        const atq_cluster_bounds cb = c.bounds[cl];
        c.cpts[pos] = atq_cluster_point{p.packed_x, p.packed_y,
                                        calc_theta(p.packed_x, p.packed_y, bounds_cx(cb),
                                                   bounds_cy(cb), fx)};
        c.cidx[pos] = cl;
    }
}

// The fallback: a stable LSD sort of the packed keys (`passes` digits of
// BITS bits) from X into Y, then the cluster stage (rs_fallback_clusters),
// by work-groups 0 .. G-1 of the last rank-pass kernel (the others exit at
// once, so these G are resident together), which run it only when
// *d.passes == 0: frames that need no fallback pay nothing for it. Each
// pass: chunk histograms, column scans, scatter, separated by grid
// barriers.
template <int BITS>
ATQ_DEV void rs_full_body(const RsDyn &d, const uint32_t *count, uint32_t cap, uint32_t passes,
                          uint32_t G, uint32_t *hist, uint32_t *totals, uint32_t *bar,
                          uint32_t *error, RsScatterLds<BITS> &L, BoundsAcc *wpart) {
    constexpr int B = 1 << BITS;
    const uint32_t n = min(*count, cap), g = blockIdx.x;
    uint32_t chunk = (n + G - 1) / G;
    chunk = (chunk + RS_ROUND - 1) / RS_ROUND * RS_ROUND;
    const uint32_t b = min(n, g * chunk), e = min(n, b + chunk);
    uint32_t gen = 0;
    for (uint32_t pass = 0; pass < passes; pass++) {
        const bool from_y = pass & 1u;
        const uint64_t *ki = from_y ? d.ky : d.kx;
        const uint32_t *vi = from_y ? d.vy : d.vx;
        uint64_t *ko = from_y ? d.kx : d.ky;
        uint32_t *vo = from_y ? d.vx : d.vy;
        const uint32_t shift = pass * BITS;
        rs_hist_body<BITS>(ki, b, e, shift, hist + (size_t)g * B, L.wc[0]);
        rs_grid_sync(bar, ++gen * G, error);
        for (uint32_t dd = g * RS_THREADS + threadIdx.x; dd < (uint32_t)B; dd += G * RS_THREADS) {
            uint32_t run = 0;
            for (uint32_t q = 0; q < G; q++) {
                const uint32_t v = hist[(size_t)q * B + dd];
                hist[(size_t)q * B + dd] = run;
                run += v;
            }
            totals[dd] = run;
        }
        rs_grid_sync(bar, ++gen * G, error);
        rs_scatter_body<BITS>(ki, vi, ko, vo, b, e, shift, hist + (size_t)g * B, totals,
                              RsFinal{}, L);
        rs_grid_sync(bar, ++gen * G, error);
    }
    if ((passes & 1u) == 0) {  // the last pass wrote X
        for (uint32_t i = g * RS_THREADS + threadIdx.x; i < n; i += G * RS_THREADS) {
            d.ky[i] = d.kx[i];
            d.vy[i] = d.vx[i];
        }
        rs_grid_sync(bar, ++gen * G, error);
    }
    rs_fallback_clusters(d, n, G, bar, gen, error, L.lds4, wpart);
}

// The last pass's extras as plain values (taking the address of the
// kernel argument would copy it to scratch memory in every lane).
ATQ_DEV RsFinal rs_final(const RsDyn &d, bool last) {
    RsFinal fin{};
    if (last) {
        fin.on = true;
        fin.shift = *d.shift;
        fin.xlat = d.xlat;
        fin.centre = d.cl.centre;
        fin.cpts = d.cl.cpts;
        fin.fx = d.cl.fx != 0;
    }
    return fin;
}

// One rank pass by the G resident work-groups of a launch: chunk
// histograms, column scans and the scatter, separated by grid barriers
// (the second pass of frames whose ranks need two digits, so that frames
// needing one pay for a single idle launch, not three).
ATQ_DEV void rs_rank_pass_body(const RsDyn &d, const uint32_t *count, uint32_t cap, uint32_t pass,
                               uint32_t G, uint32_t *hist, uint32_t *totals, uint32_t *bar,
                               uint32_t *error, RsScatterLds<RK_BITS> &L) {
    constexpr int B = 1 << RK_BITS;
    const uint64_t *ki;
    const uint32_t *vi;
    uint64_t *ko;
    uint32_t *vo;
    bool last;
    if (!rs_dyn_io(d, pass, &ki, &vi, &ko, &vo, &last)) return;
    const uint32_t n = min(*count, cap), g = blockIdx.x;
    uint32_t chunk = (n + G - 1) / G;
    chunk = (chunk + RS_ROUND - 1) / RS_ROUND * RS_ROUND;
    const uint32_t b = min(n, g * chunk), e = min(n, b + chunk);
    const uint32_t shift = pass * RK_BITS;
    uint32_t gen = 0;
    rs_hist_body<RK_BITS>(ki, b, e, shift, hist + (size_t)g * B, L.wc[0]);
    rs_grid_sync(bar, ++gen * G, error);
    for (uint32_t dd = g * RS_THREADS + threadIdx.x; dd < (uint32_t)B; dd += G * RS_THREADS) {
        uint32_t run = 0;
        for (uint32_t q = 0; q < G; q++) {
            const uint32_t v = hist[(size_t)q * B + dd];
            hist[(size_t)q * B + dd] = run;
            run += v;
        }
        totals[dd] = run;
    }
    rs_grid_sync(bar, ++gen * G, error);
    rs_scatter_body<RK_BITS>(ki, vi, ko, vo, b, e, shift, hist + (size_t)g * B, totals,
                             rs_final(d, last), L);
}

// The last launched rank pass (pass 1) also hosts the full-key fallback and
// the second rank pass: its work-groups 0 .. RS_FULL_GROUPS-1 run the
// fallback's full_passes 11-bit digits when the ranks are not used, or the
// second rank pass when the ranks need two digits (rs_rank_pass_body);
// otherwise every work-group exits at once. Launched for pass 0 it is that
// pass's scatter.
__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter_dyn(RsDyn d, const uint32_t *count, uint32_t cap, uint32_t pass,
                 uint32_t *__restrict__ hist, uint32_t *__restrict__ totals, uint32_t full_passes,
                 uint32_t *bar, uint32_t *error) {
    constexpr int B = 1 << RK_BITS;
    __shared__ RsScatterLds<RK_BITS> L;
    __shared__ BoundsAcc wpart[RS_THREADS / 64];
    if (*d.passes == 0) {
        if (full_passes && blockIdx.x < RS_FULL_GROUPS)
            rs_full_body<RK_BITS>(d, count, cap, full_passes, RS_FULL_GROUPS, hist, totals, bar,
                                  error, L, wpart);
        return;
    }
    if (pass > 0) {
        if (blockIdx.x < RS_FULL_GROUPS)
            rs_rank_pass_body(d, count, cap, pass, RS_FULL_GROUPS, hist, totals, bar, error, L);
        return;
    }
    const uint64_t *ki;
    const uint32_t *vi;
    uint64_t *ko;
    uint32_t *vo;
    bool last;
    if (!rs_dyn_io(d, pass, &ki, &vi, &ko, &vo, &last)) return;
    const uint32_t n = min(*count, cap);
    uint32_t chunk, G;
    rs_chunks(n, &chunk, &G);
    if (blockIdx.x >= G) return;
    const uint32_t b = blockIdx.x * chunk, e = min(n, b + chunk);
    rs_scatter_body<RK_BITS>(ki, vi, ko, vo, b, e, pass * RK_BITS, hist + (size_t)blockIdx.x * B,
                             totals, rs_final(d, last), L);
}

int rank_sort(hipStream_t s, uint32_t full_bits, const RsDyn &d, const uint32_t *count,
              uint32_t cap, uint32_t *hist, uint32_t *totals, uint32_t *bar, uint32_t *error) {
    constexpr uint32_t B = 1u << RK_BITS;
    const uint32_t full_passes = (full_bits + RK_BITS - 1) / RK_BITS;
    static_assert(RK_MAX_PASSES == 2, "pass 1 runs inside the last launch");
    k_rs_hist_dyn<<<RS_GROUPS, RS_THREADS, 0, s>>>(d, count, cap, 0, hist);
    k_rs_scan<<<B / 64, 64 * RS_SCAN_WAVES, 0, s>>>(hist, B, count, cap, totals, d.passes, 0);
    k_rs_scatter_dyn<<<RS_GROUPS, RS_THREADS, 0, s>>>(d, count, cap, 0, hist, totals, 0u, bar,
                                                      error);
    // pass 1 (and the fallback): RS_FULL_GROUPS work-groups, grid barriers
    k_rs_scatter_dyn<<<RS_FULL_GROUPS, RS_THREADS, 0, s>>>(d, count, cap, 1, hist, totals,
                                                           full_passes, bar, error);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

RadixPlan radix_plan(size_t npx, int extra_bits) {
    RadixPlan p;
    int lb = 1;
    while (lb < 31 && ((size_t)1 << lb) < npx) lb++;
    p.label_bits = lb;
    const int total = 2 * lb + extra_bits;
    p.passes = (total + 10) / 11;
    const int per = (total + p.passes - 1) / p.passes;
    p.digit_bits = per <= 8 ? 8 : (per <= 10 ? 10 : 11);
    return p;
}

template <int BITS>
static void rs_pass(hipStream_t s, uint32_t shift, const uint64_t *ki, const uint32_t *vi,
                    uint64_t *ko, uint32_t *vo, const uint32_t *count, uint32_t cap,
                    uint32_t *hist, uint32_t *totals) {
    constexpr uint32_t B = 1u << BITS;
    k_rs_hist<BITS><<<RS_GROUPS, RS_THREADS, 0, s>>>(ki, count, cap, shift, hist);
    k_rs_scan<<<B / 64, 64 * RS_SCAN_WAVES, 0, s>>>(hist, B, count, cap, totals);
    k_rs_scatter<BITS><<<RS_GROUPS, RS_THREADS, 0, s>>>(ki, vi, ko, vo, count, cap, shift, hist,
                                                       totals);
}

int radix_sort_packed(hipStream_t s, const RadixPlan &p, uint64_t *keys, uint32_t *vals,
                      uint64_t *keys_alt, uint32_t *vals_alt, const uint32_t *count, uint32_t cap,
                      uint32_t *hist, uint32_t *totals, bool *result_in_alt) {
    uint64_t *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    for (int pass = 0; pass < p.passes; pass++) {
        const uint32_t shift = (uint32_t)(pass * p.digit_bits);
        switch (p.digit_bits) {
        case 8: rs_pass<8>(s, shift, ki, vi, ko, vo, count, cap, hist, totals); break;
        case 10: rs_pass<10>(s, shift, ki, vi, ko, vo, count, cap, hist, totals); break;
        default: rs_pass<11>(s, shift, ki, vi, ko, vo, count, cap, hist, totals); break;
        }
        uint64_t *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
    }
    *result_in_alt = (p.passes & 1) != 0;
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_pack_keys(hipStream_t s, const uint64_t *in, uint64_t *out, const uint32_t *count,
                     uint32_t cap, uint32_t lb, bool unpack) {
    unsigned grid = min(4096u, atq_div_up(cap, 256) + 1);
    if (unpack) k_unpack_keys<<<grid, 256, 0, s>>>(in, out, count, cap, lb);
    else k_pack_keys<<<grid, 256, 0, s>>>(in, out, count, cap, lb);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
