// segsort.hpp - ascending sort of many independent segments of u64 keys.
//
// A segment is a {start, count} run of one array (clusters, peak segments).
// Every segment is cut into chunks of SEG_CHUNK keys; one 1024-thread
// work-group sorts a chunk in LDS with a stable LSD radix sort over the
// keys' high 32 bits (4 passes of 8 bits; each wave ranks its elements with
// ballots, the 16 waves' digit counts are scanned in the work-group). The
// low 32 bits ride along: callers put there a number that increases with the
// element's position (its index), so sorting by the high word alone orders
// the full 64-bit keys, which are distinct. A segment that fits one chunk is
// finished right there by the caller's Finish functor. Longer segments are
// written to keys0 and completed by merge passes in which every key finds
// its rank in the partner run by binary search; runs double each pass, so a
// segment of n keys needs seg_merge_passes(n) passes and ends in keys1 when
// that number is odd, else in keys0.
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

// LDS of a chunk sort: the keys' high and low words, and the digit
// counters (256 digits x 16 waves, digit-major).
constexpr int SEG_WAVES = SEG_NT / 64;
constexpr uint32_t SEG_WAVE_SPAN = 64 * SEG_IPT;  // elements per wave
struct SegLds {
    uint32_t hi[SEG_CHUNK];
    uint32_t lo[SEG_CHUNK];
    uint32_t cnt[256 * SEG_WAVES];
    uint32_t wsum[SEG_WAVES];
};
static_assert(256 * SEG_WAVES == 4 * SEG_NT, "four counters per thread");

// Stable LSD radix sort of L.hi/L.lo[0..m) by the high word. Wave w holds
// elements w*SEG_WAVE_SPAN + i*64 + lane (i < SEG_IPT), so ranks taken in i
// order, then by lane, follow the element order.
ATQ_DEV void seg_lds_radix(SegLds &L, uint32_t m) {
    constexpr int NW = SEG_WAVES;
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    const uint32_t wb = w * SEG_WAVE_SPAN;
    const uint64_t lt = lanemask_lt();
    for (int pass = 0; pass < 4; pass++) {
        const uint32_t shift = 8 * pass;
        uint32_t hi[SEG_IPT], lo[SEG_IPT], d[SEG_IPT], rank[SEG_IPT];
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) {
            const uint32_t e = wb + i * 64 + lane;
            hi[i] = e < m ? L.hi[e] : 0u;
            lo[i] = e < m ? L.lo[e] : 0u;
            d[i] = (hi[i] >> shift) & 255u;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) L.cnt[4 * t + k] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) {
            if (wb + i * 64 >= m) break;  // the rest of this wave is empty
            const bool valid = wb + i * 64 + lane < m;
            uint64_t mk = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; bit++) {
                const uint64_t bb = __ballot((d[i] >> bit) & 1u);
                mk &= ((d[i] >> bit) & 1u) ? bb : ~bb;
            }
            const int leader = valid ? __ffsll((unsigned long long)mk) - 1 : (int)lane;
            uint32_t before = 0;
            if (valid && (int)lane == leader) {
                before = L.cnt[d[i] * NW + w];
                L.cnt[d[i] * NW + w] = before + (uint32_t)__popcll(mk);
            }
            before = __shfl(before, leader, 64);
            rank[i] = before + (uint32_t)__popcll(mk & lt);
        }
        __syncthreads();
        uint32_t c4[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            c4[k] = L.cnt[4 * t + k];
            sum += c4[k];
        }
        uint32_t all;
        uint32_t run = block_exclusive_scan<uint32_t, SEG_NT>(sum, L.wsum, &all);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            L.cnt[4 * t + k] = run;
            run += c4[k];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) {
            if (wb + i * 64 + lane >= m) continue;
            const uint32_t pos = L.cnt[d[i] * NW + w] + rank[i];
            L.hi[pos] = hi[i];
            L.lo[pos] = lo[i];
        }
        __syncthreads();
    }
}

// KeyGen: uint64_t operator()(uint32_t seg, uint32_t start, uint32_t j) const
// (key of element j of the segment at `start`). Ext: {start, count}.
// Finish: void operator()(uint32_t seg, uint32_t start, uint32_t n,
// const uint64_t (&v)[SEG_IPT], uint64_t *lds) const, called by every thread
// of the work-group with a whole sorted segment of n <= SEG_CHUNK keys in
// registers (key e0 + i in v[i], e0 = threadIdx.x * SEG_IPT) and lds, 2048
// u64 of scratch.
// Keys move between global memory and LDS coalesced.
template <typename KeyGen, typename Finish, typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_chunks(KeyGen gen, Finish fin, const Ext *__restrict__ segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, uint32_t max_chunks, uint64_t *__restrict__ keys0) {
    __shared__ SegLds L;
    const uint32_t nseg = min(*n_segs_dev, cap_segs);
    const uint64_t items = (uint64_t)nseg * max_chunks;
    const uint32_t t = threadIdx.x, e0 = t * SEG_IPT;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t sg = (uint32_t)(it / max_chunks), ch = (uint32_t)(it % max_chunks);
        const uint32_t start = segs[sg].start, n = segs[sg].count;
        const uint32_t b = ch * SEG_CHUNK;
        if (b >= n) continue;
        const uint32_t m = min(SEG_CHUNK, n - b);
        for (uint32_t j = t; j < m; j += SEG_NT) {
            const uint64_t k = gen(sg, start, b + j);
            L.hi[j] = (uint32_t)(k >> 32);
            L.lo[j] = (uint32_t)k;
        }
        __syncthreads();
        seg_lds_radix(L, m);
        if (n <= SEG_CHUNK) {
            uint64_t v[SEG_IPT];
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++)
                v[i] = e0 + i < m ? ((uint64_t)L.hi[e0 + i] << 32) | L.lo[e0 + i] : SEG_PAD;
            __syncthreads();
            fin(sg, start, n, v, reinterpret_cast<uint64_t *>(L.cnt));
            __syncthreads();
            continue;
        }
        for (uint32_t j = t; j < m; j += SEG_NT)
            keys0[start + b + j] = ((uint64_t)L.hi[j] << 32) | L.lo[j];
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
