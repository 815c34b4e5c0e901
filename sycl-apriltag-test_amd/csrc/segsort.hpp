// segsort.hpp - ascending sort of many independent segments of u64 keys.
//
// A segment is a {start, count} run of one array (clusters, peak segments).
// Every segment is cut into chunks of SEG_CHUNK keys; one 1024-thread
// work-group sorts a chunk with a bitonic network whose keys stay in
// registers (4 per thread): exchanges within a thread are register moves,
// within a wave lane shuffles, and only the exchanges across the 16 waves go
// through LDS. A segment that fits one chunk is finished right there, from
// the registers, by the caller's Finish functor. Longer segments are written
// to keys0 and completed by merge passes in which every key finds its rank in
// the partner run by binary search; runs double each pass, so a segment of n
// keys needs seg_merge_passes(n) passes and ends in keys1 when that number is
// odd, else in keys0. Keys must be distinct (callers put the element's
// position in the low bits), so the result is a stable sort.
#pragma once
#include "common.hpp"

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

ATQ_DEV void seg_cas(uint64_t &a, uint64_t &b, bool ascending) {
    const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = ascending ? lo : hi;
    b = ascending ? hi : lo;
}

template <int J>
ATQ_DEV void seg_regs_step(uint64_t (&v)[SEG_IPT], uint32_t k, uint32_t t) {
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) {
        if (i & J) continue;
        const uint32_t e = t * SEG_IPT + i;
        seg_cas(v[i], v[i + J], (e & k) == 0);
    }
}

ATQ_DEV uint32_t seg_lds_index(uint32_t e) { return e + (e >> 4); }

// Bitonic sort (ascending) of the first P keys (P a power of two <= SEG_CHUNK)
// held blocked in registers: thread t holds keys t*SEG_IPT .. t*SEG_IPT+3.
// Keys past P must be SEG_PAD. All SEG_NT threads call it with the same P;
// lds holds SEG_CHUNK + SEG_CHUNK/16 keys.
ATQ_DEV void seg_block_sort(uint64_t (&v)[SEG_IPT], uint32_t P, uint64_t *lds) {
    const uint32_t t = threadIdx.x, lane = t & 63, e0 = t * SEG_IPT;
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j < (uint32_t)SEG_IPT) {
                if (j == 2) seg_regs_step<2>(v, k, t);
                else seg_regs_step<1>(v, k, t);
            } else if (j < (uint32_t)SEG_IPT * 64) {
                // partner = same register of lane ^ m; the element keeps the
                // minimum iff (it is the lower of the pair) == (ascending)
                const uint32_t m = j / SEG_IPT;
                const bool keep_min = ((lane & m) == 0) == ((e0 & k) == 0);
#pragma unroll
                for (int i = 0; i < SEG_IPT; i++) {
                    const uint64_t o = __shfl_xor(v[i], (int)m, 64);
                    v[i] = keep_min ? (o < v[i] ? o : v[i]) : (o > v[i] ? o : v[i]);
                }
            } else {
#pragma unroll
                for (int i = 0; i < SEG_IPT; i++) lds[seg_lds_index(e0 + i)] = v[i];
                __syncthreads();
                const bool keep_min = ((e0 & j) == 0) == ((e0 & k) == 0);
#pragma unroll
                for (int i = 0; i < SEG_IPT; i++) {
                    const uint64_t o = lds[seg_lds_index((e0 + i) ^ j)];
                    v[i] = keep_min ? (o < v[i] ? o : v[i]) : (o > v[i] ? o : v[i]);
                }
                __syncthreads();
            }
        }
    }
}

// KeyGen: uint64_t operator()(uint32_t seg, uint32_t start, uint32_t j) const
// (key of element j of the segment at `start`). Ext: {start, count}.
// Finish: void operator()(uint32_t seg, uint32_t start, uint32_t n,
// const uint64_t (&v)[SEG_IPT], uint64_t *lds) const, called by every thread
// of the work-group with a whole sorted segment of n <= SEG_CHUNK keys in
// registers (key e0 + i in v[i], e0 = threadIdx.x * SEG_IPT; lds is free).
// Keys move between global memory and registers through LDS, so that both
// the global loads and the stores are coalesced.
template <typename KeyGen, typename Finish, typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_chunks(KeyGen gen, Finish fin, const Ext *__restrict__ segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, uint32_t max_chunks, uint64_t *__restrict__ keys0) {
    __shared__ uint64_t lds[SEG_CHUNK + SEG_CHUNK / 16];
    const uint32_t nseg = min(*n_segs_dev, cap_segs);
    const uint64_t items = (uint64_t)nseg * max_chunks;
    const uint32_t t = threadIdx.x, e0 = t * SEG_IPT;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t sg = (uint32_t)(it / max_chunks), ch = (uint32_t)(it % max_chunks);
        const uint32_t start = segs[sg].start, n = segs[sg].count;
        const uint32_t b = ch * SEG_CHUNK;
        if (b >= n) continue;
        const uint32_t m = min(SEG_CHUNK, n - b);
        uint32_t P = 1;
        while (P < m) P <<= 1;
        for (uint32_t j = t; j < m; j += SEG_NT) lds[seg_lds_index(j)] = gen(sg, start, b + j);
        __syncthreads();
        uint64_t v[SEG_IPT];
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) v[i] = e0 + i < m ? lds[seg_lds_index(e0 + i)] : SEG_PAD;
        __syncthreads();
        seg_block_sort(v, P, lds);
        if (n <= SEG_CHUNK) {
            fin(sg, start, n, v, lds);
            __syncthreads();
            continue;
        }
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++)
            if (e0 + i < m) lds[seg_lds_index(e0 + i)] = v[i];
        __syncthreads();
        for (uint32_t j = t; j < m; j += SEG_NT) keys0[start + b + j] = lds[seg_lds_index(j)];
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
