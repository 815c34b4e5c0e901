// segsort.hpp - ascending sort of many independent segments of u64 keys.
//
// A segment is a {start, count} run of one array (clusters, peak segments).
// Every segment is cut into chunks of SEG_CHUNK keys; one 1024-thread
// work-group sorts a chunk in LDS: each thread sorts 4 keys in registers, then
// runs are merged pairwise (merge path: every thread finds where its 4
// outputs start by binary search on its diagonal and merges them), 10
// levels for a full chunk. Keys must be distinct (callers put the element's
// position in the low bits), so the result is a stable sort. A segment that
// fits one chunk is finished right there by the caller's Finish functor.
// Longer segments are written to keys0 and completed by merge passes in
// which every key finds its rank in the partner run by binary search; runs
// double each pass, so a segment of n keys needs seg_merge_passes(n) passes
// and ends in keys1 when that number is odd, else in keys0.
#pragma once
#include "scan.hpp"

namespace atq {

constexpr int SEG_NT = 1024;
constexpr int SEG_IPT = 4;
constexpr uint32_t SEG_CHUNK = SEG_NT * SEG_IPT;
constexpr uint64_t SEG_PAD = ~0ull;  // sorts after every real key

ATQ_DEV uint32_t seg_merge_passes(uint32_t n) {
    uint32_t q = 0;
    for (uint32_t r = SEG_CHUNK; r < n; r <<= 1) q++;
    return q;
}

inline uint32_t seg_max_chunks(uint32_t max_seg) {
    return max_seg ? (max_seg + SEG_CHUNK - 1) / SEG_CHUNK : 1;
}

// Inclusive scan of six int64 sums over the SEG_NT threads of a work-group
// (lds: at least 6 * SEG_NT / 64 + 6 entries); tot gets the work-group
// totals. Wave 0 scans the per-wave totals in place.
ATQ_DEV void seg_block_scan6(const int64_t (&v)[6], int64_t (&incl)[6], int64_t (&tot)[6],
                             int64_t *lds) {
    constexpr int NW = SEG_NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        int64_t x = v[f];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        incl[f] = x;
        if (lane == 63) lds[f * NW + wave] = x;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const int64_t t = lane < NW ? lds[f * NW + lane] : 0;
            int64_t x = t;
#pragma unroll
            for (int o = 1; o < NW; o <<= 1) {
                const int64_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane < NW) lds[f * NW + lane] = x - t;
            if (lane == NW - 1) lds[6 * NW + f] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 6; f++) {
        incl[f] += lds[f * NW + wave];
        tot[f] = lds[6 * NW + f];
    }
    __syncthreads();
}

ATQ_DEV void seg_cas(uint64_t &a, uint64_t &b) {
    const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Sorts a[0..m) (m <= SEG_CHUNK; a[m..SEG_CHUNK) must hold SEG_PAD) by merge
// path; b is a second SEG_CHUNK buffer. The result is in a. All SEG_NT
// threads call it.
ATQ_DEV void seg_merge_sort(uint64_t *a, uint64_t *b, uint32_t m) {
    static_assert(SEG_IPT == 4, "four keys per thread");
    const uint32_t t = threadIdx.x, d0 = t * SEG_IPT;
    uint32_t P = SEG_IPT;
    while (P < m) P <<= 1;
    uint64_t v[SEG_IPT];
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) v[i] = a[d0 + i];
    seg_cas(v[0], v[1]);
    seg_cas(v[2], v[3]);
    seg_cas(v[0], v[2]);
    seg_cas(v[1], v[3]);
    seg_cas(v[1], v[2]);
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) a[d0 + i] = v[i];
    __syncthreads();
    uint64_t *src = a, *dst = b;
    for (uint32_t r = SEG_IPT; r < P; r <<= 1) {
        if (d0 < P) {
            const uint32_t base = d0 / (2 * r) * (2 * r), d = d0 - base;
            const uint64_t *A = src + base, *B = src + base + r;
            // i = number of A keys among the first d outputs
            uint32_t lo = d > r ? d - r : 0, hi = d < r ? d : r;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (A[mid] < B[d - mid - 1]) lo = mid + 1;
                else hi = mid;
            }
            uint32_t i = lo, j = d - lo;
#pragma unroll
            for (int k = 0; k < SEG_IPT; k++) {
                const bool take_a = j >= r || (i < r && A[i] < B[j]);
                dst[base + d + k] = take_a ? A[i] : B[j];
                if (take_a) i++;
                else j++;
            }
        }
        __syncthreads();
        uint64_t *tmp = src;
        src = dst;
        dst = tmp;
    }
    if (src != a) {
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) a[d0 + i] = src[d0 + i];
        __syncthreads();
    }
}

// KeyGen: uint64_t operator()(uint32_t seg, uint32_t start, uint32_t j) const
// (key of element j of the segment at `start`). Ext: {start, count}.
// Finish: void operator()(uint32_t seg, uint32_t start, uint32_t n,
// const uint64_t (&v)[SEG_IPT], uint64_t *lds) const, called by every thread
// of the work-group with a whole sorted segment of n <= SEG_CHUNK keys in
// registers (key e0 + i in v[i], e0 = threadIdx.x * SEG_IPT) and lds,
// 2 * SEG_CHUNK u64 of scratch: the second half is free on entry, the first
// half (where the keys were read from) once a barrier has passed. Keys move
// between global memory and LDS coalesced.
template <typename KeyGen, typename Finish, typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_chunks(KeyGen gen, Finish fin, const Ext *__restrict__ segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, uint32_t max_chunks, uint64_t *__restrict__ keys0) {
    __shared__ uint64_t kab[2 * SEG_CHUNK];
    uint64_t *const ka = kab, *const kb = kab + SEG_CHUNK;
    const uint32_t nseg = min(*n_segs_dev, cap_segs);
    const uint64_t items = (uint64_t)nseg * max_chunks;
    const uint32_t t = threadIdx.x, e0 = t * SEG_IPT;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t sg = (uint32_t)(it / max_chunks), ch = (uint32_t)(it % max_chunks);
        const uint32_t start = segs[sg].start, n = segs[sg].count;
        const uint32_t b = ch * SEG_CHUNK;
        if (b >= n) continue;
        const uint32_t m = min(SEG_CHUNK, n - b);
        for (uint32_t j = t; j < SEG_CHUNK; j += SEG_NT)
            ka[j] = j < m ? gen(sg, start, b + j) : SEG_PAD;
        __syncthreads();
        seg_merge_sort(ka, kb, m);
        if (n <= SEG_CHUNK) {
            uint64_t v[SEG_IPT];
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++) v[i] = ka[e0 + i];
            fin(sg, start, n, v, kab);
            __syncthreads();
            continue;
        }
        for (uint32_t j = t; j < m; j += SEG_NT) keys0[start + b + j] = ka[j];
        __syncthreads();
    }
}

ATQ_DEV uint32_t seg_lower_bound(const uint64_t *a, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One merge pass, one thread per element: sorted runs of length `run` pair
// up into runs of 2*run, for every segment longer than `run`. seg_of[x] is
// the segment of element x.
template <typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_merge(const Ext *__restrict__ segs, const uint32_t *__restrict__ seg_of,
            const uint32_t *n_items_dev, uint32_t cap_items, uint32_t run,
            const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout) {
    const uint32_t N = min(*n_items_dev, cap_items);
    for (uint32_t x = blockIdx.x * SEG_NT + threadIdx.x; x < N; x += gridDim.x * SEG_NT) {
        const Ext e = segs[seg_of[x]];
        if (e.count <= run) continue;
        const uint32_t n = e.count, j = x - e.start;
        const uint64_t *c = kin + e.start;
        const uint32_t base = j / (2 * run) * (2 * run);
        const uint64_t key = c[j];
        uint32_t pos;
        if (j - base < run) {  // left run: add the right run's smaller keys
            const uint32_t rb = min(base + run, n), re = min(base + 2 * run, n);
            pos = j + seg_lower_bound(c + rb, re - rb, key);
        } else {               // right run: add the left run's smaller keys
            pos = j - run + seg_lower_bound(c + base, run, key);
        }
        kout[e.start + pos] = key;
    }
}

// Sorts every segment. The segments tile elements [0, *n_items_dev) and
// seg_of maps each element to its segment; max_seg bounds the segment length
// (longer segments are left partly sorted: callers bound it by construction).
template <typename KeyGen, typename Finish, typename Ext>
int seg_sort(hipStream_t s, KeyGen gen, Finish fin, const Ext *segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, const uint32_t *seg_of, const uint32_t *n_items_dev,
             uint32_t cap_items, uint32_t max_seg, uint64_t *keys0, uint64_t *keys1) {
    const uint32_t max_chunks = seg_max_chunks(max_seg);
    k_seg_chunks<KeyGen, Finish, Ext><<<1024, SEG_NT, 0, s>>>(gen, fin, segs, n_segs_dev, cap_segs,
                                                               max_chunks, keys0);
    const unsigned grid = min(8192u, (cap_items + SEG_NT - 1) / SEG_NT + 1);
    int pass = 0;
    for (uint32_t run = SEG_CHUNK; run < max_seg; run <<= 1, pass++)
        k_seg_merge<Ext><<<grid, SEG_NT, 0, s>>>(segs, seg_of, n_items_dev, cap_items, run,
                                                  (pass & 1) ? keys1 : keys0,
                                                  (pass & 1) ? keys0 : keys1);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
