// segsort.hpp - ascending sort of many independent segments of u64 keys.
//
// A segment is a {start, count} run of one array (clusters, peak segments).
// Every segment is cut into chunks of SEG_CHUNK keys; one 1024-thread
// work-group sorts a chunk in LDS: each thread sorts 4 keys in registers, then
// runs are merged pairwise (merge path: every thread finds where its 4
// outputs start by binary search on its diagonal and merges them), 10
// levels for a full chunk. Keys must be distinct (callers put the element's
// position in the low bits), so the result is a stable sort. A key generator
// whose low word is exactly the element's index in its segment
// (KeyGen::position_low) has its chunks sorted on the high words alone, with
// the position as a 16-bit payload and a stable merge (ties keep position
// order, which is the order of the low words): half the LDS traffic of the
// 64-bit merge for the same result. A segment that
// fits one chunk is finished right there by the caller's Finish functor.
// Longer segments are written to keys0 and named in a device-side list of
// long segments. One work-group then merges each SEG_LDS_KEYS block of a
// listed segment in LDS (keys0 -> keys1); segments longer than that take
// global merge passes, in which every key finds its rank in the partner run
// by binary search, alternating keys1 -> keys0 -> keys1 ...;
// seg_sorted_keys(n, ...) names the buffer a segment of n keys ends in.
#pragma once
#include <type_traits>

#include "scan.hpp"

namespace atq {

constexpr int SEG_NT = 1024;
constexpr int SEG_IPT = 4;
constexpr uint32_t SEG_CHUNK = SEG_NT * SEG_IPT;
constexpr uint64_t SEG_PAD = ~0ull;  // sorts after every real key

constexpr uint32_t SEG_LDS_KEYS = 16384;  // keys one work-group merges in LDS
constexpr uint32_t SEG_LDS_IPT = SEG_LDS_KEYS / SEG_NT;
constexpr uint32_t SEG_MERGE_NT = 256;    // global passes: keys per work-group item
constexpr uint32_t SEG_MERGE_ITEM = 4 * SEG_MERGE_NT;

// Whether the long segments of a sort bounded by max_seg are merged in LDS
// first (SEG_CHUNK runs into SEG_LDS_KEYS blocks, one work-group each) or
// by global merge passes from SEG_CHUNK on: the LDS merge pays when many
// long segments share the device (frames whose clusters can pass
// SEG_LDS_KEYS points, or several frames at once), the global passes when
// few do (they spread one segment over many work-groups, for latency).
__host__ __device__ inline bool seg_lds_merge(uint32_t max_seg) { return max_seg > SEG_LDS_KEYS; }

// Global merge passes a segment of n > SEG_CHUNK keys takes after the LDS
// merge (runs of SEG_LDS_KEYS up) or without it (runs of SEG_CHUNK up).
ATQ_DEV uint32_t seg_global_passes(uint32_t n, bool lds) {
    uint32_t q = 0;
    for (uint32_t r = lds ? SEG_LDS_KEYS : SEG_CHUNK; r < n; r <<= 1) q++;
    return q;
}

// Where the sorted keys of a segment of n keys end (lds = seg_lds_merge of
// the sort's bound): keys1 when it fits one chunk (the Finish functor
// writes them there); otherwise after an odd number of hops keys0 -> keys1
// -> keys0 ... keys1, else keys0 (the LDS merge is one hop).
template <typename T>
ATQ_DEV T *seg_sorted_keys(uint32_t n, T *keys0, T *keys1, bool lds) {
    if (n <= SEG_CHUNK) return keys1;
    const uint32_t hops = seg_global_passes(n, lds) + (lds ? 1u : 0u);
    return (hops & 1) ? keys1 : keys0;
}

inline uint32_t seg_max_chunks(uint32_t max_seg) {
    return max_seg ? (max_seg + SEG_CHUNK - 1) / SEG_CHUNK : 1;
}

// Entries a long-segment list needs for segments tiling cap_items elements.
inline uint32_t seg_long_cap(uint32_t cap_items) { return cap_items / (SEG_CHUNK + 1) + 1; }

// The segments longer than SEG_CHUNK, in no particular order; n is zero
// before the sort that fills it.
struct SegLong {
    uint32_t *n;
    uint32_t *list;
    uint32_t cap;
};

// Inclusive scan of six int64 sums over the SEG_NT threads of a work-group
// (lds: at least 6 * SEG_NT / 64 + 6 entries); tot gets the work-group
// totals. Wave 0 scans the per-wave totals in place.
ATQ_DEV void seg_block_scan6(const int64_t (&v)[6], int64_t (&incl)[6], int64_t (&tot)[6],
                             int64_t *lds) {
    constexpr int NW = SEG_NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    static_assert(NW == 16, "the per-wave totals fill one DPP row");
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int64_t x = wave_inclusive_sum(v[f]);
        incl[f] = x;
        if (lane == 63) lds[f * NW + wave] = x;
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const int64_t t = lane < NW ? lds[f * NW + lane] : 0;
            const int64_t x = row_inclusive_sum(t);
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

// LDS key buffers carry one pad word per 32 keys: merge-path threads read
// keys a power-of-two stride apart, which would otherwise fall into a few
// LDS banks.
ATQ_DEV uint32_t seg_ix(uint32_t i) { return i + (i >> 5); }
constexpr uint32_t seg_padded(uint32_t n) { return n + n / 32; }

// Merges the sorted runs src[base, base + r) and src[base + r, base + 2r)
// (padded indices) from output diagonal d: the first `per` outputs from d
// on, into v. Each step loads only the key of the run that advanced.
template <int N>
ATQ_DEV void seg_merge_at(const uint64_t *src, uint32_t base, uint32_t r, uint32_t d,
                          uint32_t per, uint64_t (&v)[N]) {
    // i = number of left-run keys among the first d outputs
    uint32_t lo = d > r ? d - r : 0, hi = d < r ? d : r;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (src[seg_ix(base + mid)] < src[seg_ix(base + r + d - mid - 1)]) lo = mid + 1;
        else hi = mid;
    }
    uint32_t i = lo, j = d - lo;
    uint64_t av = i < r ? src[seg_ix(base + i)] : SEG_PAD;
    uint64_t bv = j < r ? src[seg_ix(base + r + j)] : SEG_PAD;
#pragma unroll
    for (int k = 0; k < N; k++) {
        if ((uint32_t)k >= per) break;
        const bool take_a = j >= r || (i < r && av < bv);
        v[k] = take_a ? av : bv;
        if (take_a) i++;
        else j++;
        if (k + 1 < N) {
            const bool more = take_a ? i < r : j < r;
            const uint64_t nv = more ? src[seg_ix(take_a ? base + i : base + r + j)] : SEG_PAD;
            if (take_a) av = nv;
            else bv = nv;
        }
    }
}

// Sorts a[0..m) (m <= SEG_CHUNK; a[m..SEG_CHUNK) must hold SEG_PAD; indices
// padded by seg_ix) by merge path; b is a second buffer of the same size. The
// result is in a. All SEG_NT threads call it.
ATQ_DEV void seg_merge_sort(uint64_t *a, uint64_t *b, uint32_t m) {
    static_assert(SEG_IPT == 4, "four keys per thread");
    const uint32_t t = threadIdx.x, d0 = t * SEG_IPT;
    uint32_t P = SEG_IPT;
    while (P < m) P <<= 1;
    uint64_t v[SEG_IPT];
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) v[i] = a[seg_ix(d0 + i)];
    seg_cas(v[0], v[1]);
    seg_cas(v[2], v[3]);
    seg_cas(v[0], v[2]);
    seg_cas(v[1], v[3]);
    seg_cas(v[1], v[2]);
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) a[seg_ix(d0 + i)] = v[i];
    __syncthreads();
    uint64_t *src = a, *dst = b;
    for (uint32_t r = SEG_IPT; r < P; r <<= 1) {
        if (d0 < P) {
            const uint32_t base = d0 / (2 * r) * (2 * r);
            seg_merge_at(src, base, r, d0 - base, SEG_IPT, v);
#pragma unroll
            for (int k = 0; k < SEG_IPT; k++) dst[seg_ix(d0 + k)] = v[k];
        }
        __syncthreads();
        uint64_t *tmp = src;
        src = dst;
        dst = tmp;
    }
    if (src != a) {
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) a[seg_ix(d0 + i)] = src[seg_ix(d0 + i)];
        __syncthreads();
    }
}

// Orders a wave's LDS accesses before and after it (a wave's LDS operations
// complete in issue order, so data one lane wrote is there for the others).
ATQ_DEV void seg_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Positions one wave holds in the blocked layout of the chunk sorts: merges
// of runs up to half of it read only what the same wave wrote.
constexpr uint32_t SEG_WAVE_KEYS = 64 * SEG_IPT;

// KeyGen::position_low, false when absent.
template <typename T, typename = void>
struct SegPositionLow : std::false_type {};
template <typename T>
struct SegPositionLow<T, std::void_t<decltype(T::position_low)>>
    : std::integral_constant<bool, T::position_low> {};

// Stable sort of the chunk [b, b + m) of a position-low segment: g holds the
// keys of chunk positions t + SEG_NT * i (striped; SEG_PAD past m). Sorts
// the high words with the chunk position (key - b) as payload in the u32/u16
// buffers K0, K1 (SEG_CHUNK each) and J0, J1, and returns the sorted 64-bit
// keys of positions e0 + i (e0 = threadIdx.x * SEG_IPT) in v. All SEG_NT
// threads call it; the buffers are free again once a barrier has passed.
ATQ_DEV void seg_sort_positions(const uint64_t (&g)[SEG_IPT], uint32_t m, uint32_t b,
                                uint32_t *K0, uint32_t *K1, uint16_t *J0, uint16_t *J1,
                                uint64_t (&v)[SEG_IPT]) {
    const uint32_t t = threadIdx.x, d0 = t * SEG_IPT;
    uint32_t P = SEG_IPT;
    while (P < m) P <<= 1;
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) {
        const uint32_t j = t + SEG_NT * i;
        K1[j] = j < m ? (uint32_t)(g[i] >> 32) : 0xFFFFFFFFu;  // padding sorts last (stable)
        J1[j] = (uint16_t)j;
    }
    __syncthreads();
    // four keys per thread in registers, compared with their positions
    uint64_t r[SEG_IPT];
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++) r[i] = ((uint64_t)K1[d0 + i] << 32) | J1[d0 + i];
    seg_cas(r[0], r[1]);
    seg_cas(r[2], r[3]);
    seg_cas(r[0], r[2]);
    seg_cas(r[1], r[3]);
    seg_cas(r[1], r[2]);
    const bool act = d0 < P;
    if (act)
#pragma unroll
        for (int i = 0; i < SEG_IPT; i++) {
            K0[d0 + i] = (uint32_t)(r[i] >> 32);
            J0[d0 + i] = (uint16_t)r[i];
        }
    seg_wave_sync();  // the merges of small runs stay within the wave
    uint32_t *Ks = K0, *Kd = K1;
    uint16_t *Js = J0, *Jd = J1;
    for (uint32_t run = SEG_IPT; run < P; run <<= 1) {
        if (act) {
            // merge path: i = keys of the left run among the first d outputs;
            // a left key goes first on a tie
            const uint32_t base = d0 / (2 * run) * (2 * run), d = d0 - base;
            uint32_t lo = d > run ? d - run : 0, hi = d < run ? d : run;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (Ks[base + mid] <= Ks[base + run + d - mid - 1]) lo = mid + 1;
                else hi = mid;
            }
            uint32_t i = lo, j = d - lo;
            uint32_t ak = i < run ? Ks[base + i] : 0xFFFFFFFFu;
            uint32_t bk = j < run ? Ks[base + run + j] : 0xFFFFFFFFu;
            uint32_t ok[SEG_IPT];
            uint16_t oj[SEG_IPT];
#pragma unroll
            for (int k = 0; k < SEG_IPT; k++) {
                const bool ta = j >= run || (i < run && ak <= bk);
                ok[k] = ta ? ak : bk;
                oj[k] = Js[ta ? base + i : base + run + j];
                if (ta) i++;
                else j++;
                if (k + 1 < SEG_IPT) {
                    const bool more = ta ? i < run : j < run;
                    const uint32_t nk = more ? Ks[ta ? base + i : base + run + j] : 0xFFFFFFFFu;
                    if (ta) ak = nk;
                    else bk = nk;
                }
            }
#pragma unroll
            for (int k = 0; k < SEG_IPT; k++) {
                Kd[d0 + k] = ok[k];
                Jd[d0 + k] = oj[k];
            }
        }
        if (4 * run <= SEG_WAVE_KEYS) seg_wave_sync();  // the next merge is the wave's own
        else __syncthreads();
        uint32_t *tk = Ks;
        Ks = Kd;
        Kd = tk;
        uint16_t *tj = Js;
        Js = Jd;
        Jd = tj;
    }
#pragma unroll
    for (int i = 0; i < SEG_IPT; i++)
        v[i] = d0 + i < m ? ((uint64_t)Ks[d0 + i] << 32) | (b + Js[d0 + i]) : SEG_PAD;
}

// KeyGen: uint64_t operator()(uint32_t seg, uint32_t start, uint32_t j) const
// (key of element j of the segment at `start`). Ext: {start, count}.
// Finish: void operator()(uint32_t seg, uint32_t start, uint32_t n,
// const uint64_t (&v)[SEG_IPT], uint64_t *lds) const, called by every thread
// of the work-group with a whole sorted segment of n <= SEG_CHUNK keys in
// registers (key e0 + i in v[i], e0 = threadIdx.x * SEG_IPT) and lds,
// 2 * seg_padded(SEG_CHUNK) u64 of scratch (none for a position-low
// KeyGen): the second half is free on entry, the first half (where the keys
// were read from) once a barrier has passed. Keys move
// between global memory and LDS coalesced.
template <typename KeyGen, typename Finish, typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_chunks(KeyGen gen, Finish fin, const Ext *__restrict__ segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, uint32_t max_chunks, uint64_t *__restrict__ keys0, SegLong lng) {
    constexpr bool POS = SegPositionLow<KeyGen>::value;
    // the position sort needs its u32/u16 buffers only (48 KiB: three
    // work-groups per CU); the 64-bit merge and its Finish two padded chunks
    constexpr uint32_t KAB = POS ? SEG_CHUNK * 3 / 2 : 2 * seg_padded(SEG_CHUNK);
    __shared__ uint64_t kab[KAB];
    static_assert(2 * SEG_CHUNK * (sizeof(uint32_t) + sizeof(uint16_t)) <= sizeof(kab), "");
    uint64_t *const ka = kab, *const kb = kab + seg_padded(SEG_CHUNK);
    uint32_t *const K = reinterpret_cast<uint32_t *>(kab);
    uint16_t *const J = reinterpret_cast<uint16_t *>(K + 2 * SEG_CHUNK);
    const uint32_t nseg = min(*n_segs_dev, cap_segs);
    const uint64_t items = (uint64_t)nseg * max_chunks;
    const uint32_t t = threadIdx.x, e0 = t * SEG_IPT;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {
        // chunk-major: every segment's first chunk, then every second chunk, ...,
        // so that the chunks that exist spread evenly over the work-groups
        const uint32_t sg = (uint32_t)(it % nseg), ch = (uint32_t)(it / nseg);
        const uint32_t start = segs[sg].start, n = segs[sg].count;
        const uint32_t b = ch * SEG_CHUNK;
        if (b >= n) continue;
        const uint32_t m = min(SEG_CHUNK, n - b);
        uint64_t v[SEG_IPT];
        {  // every load of a thread in flight before its first LDS store
            uint64_t g[SEG_IPT];
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++) {
                const uint32_t j = t + SEG_NT * i;
                g[i] = j < m ? gen(sg, start, b + j) : SEG_PAD;
            }
            if constexpr (POS) {
                seg_sort_positions(g, m, b, K, K + SEG_CHUNK, J, J + SEG_CHUNK, v);
            } else {
#pragma unroll
                for (int i = 0; i < SEG_IPT; i++) ka[seg_ix(t + SEG_NT * i)] = g[i];
                __syncthreads();
                seg_merge_sort(ka, kb, m);
#pragma unroll
                for (int i = 0; i < SEG_IPT; i++) v[i] = ka[seg_ix(e0 + i)];
            }
        }
        if (n <= SEG_CHUNK) {
            fin(sg, start, n, v, kab);
            __syncthreads();
            continue;
        }
        if (ch == 0 && t == 0) {
            const uint32_t slot = atomicAdd(lng.n, 1u);
            if (slot < lng.cap) lng.list[slot] = sg;
        }
        if constexpr (POS) {
#pragma unroll
            for (int i = 0; i < SEG_IPT; i++)
                if (e0 + i < m) keys0[start + b + e0 + i] = v[i];
        } else {
            for (uint32_t j = t; j < m; j += SEG_NT) keys0[start + b + j] = ka[seg_ix(j)];
        }
        __syncthreads();
    }
}

// Merges the SEG_CHUNK runs of every SEG_LDS_KEYS block of the listed long
// segments of more than min_n keys in LDS: kin (runs of SEG_CHUNK) -> kout
// (runs of SEG_LDS_KEYS, or the whole segment). Each level, every thread
// finds the start of its outputs on its merge-path diagonal and merges them
// into registers.
template <typename Ext>
__global__ void __launch_bounds__(SEG_NT)
k_seg_merge_lds(const Ext *__restrict__ segs, SegLong lng, uint32_t max_blocks,
                const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout, uint32_t min_n) {
    __shared__ uint64_t buf[seg_padded(SEG_LDS_KEYS)];
    const uint32_t nl = min(*lng.n, lng.cap);
    const uint64_t items = (uint64_t)nl * max_blocks;
    const uint32_t t = threadIdx.x;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {  // block-major
        const Ext e = segs[lng.list[it % nl]];
        const uint32_t b = (uint32_t)(it / nl) * SEG_LDS_KEYS;
        if (b >= e.count || e.count <= min_n) continue;
        const uint32_t m = min(SEG_LDS_KEYS, e.count - b);
        uint32_t P = SEG_CHUNK;
        while (P < m) P <<= 1;
        const uint32_t per = P / SEG_NT, d0 = t * per;
        uint64_t v[SEG_LDS_IPT];
#pragma unroll
        for (uint32_t k = 0; k < SEG_LDS_IPT; k++) {  // all loads before the first LDS store
            const uint32_t j = t + SEG_NT * k;
            v[k] = j < m ? kin[e.start + b + j] : SEG_PAD;
        }
#pragma unroll
        for (uint32_t k = 0; k < SEG_LDS_IPT; k++)
            if (k < per) buf[seg_ix(t + SEG_NT * k)] = v[k];
        __syncthreads();
        // threads whose outputs all lie past the segment's keys leave their
        // pads in place: a merge of pads yields the same pads
        const bool live = d0 < m;
        for (uint32_t run = SEG_CHUNK; run < P; run <<= 1) {
            const uint32_t base = d0 / (2 * run) * (2 * run);
            if (live) seg_merge_at(buf, base, run, d0 - base, per, v);
            __syncthreads();
            if (live)
#pragma unroll
                for (uint32_t k = 0; k < SEG_LDS_IPT; k++)
                    if (k < per) buf[seg_ix(d0 + k)] = v[k];
            __syncthreads();
        }
#pragma unroll
        for (uint32_t k = 0; k < SEG_LDS_IPT; k++) {
            const uint32_t j = t + SEG_NT * k;
            if (j < m) v[k] = buf[seg_ix(j)];
        }
#pragma unroll
        for (uint32_t k = 0; k < SEG_LDS_IPT; k++) {
            const uint32_t j = t + SEG_NT * k;
            if (j < m) kout[e.start + b + j] = v[k];
        }
        __syncthreads();
    }
}

// One global merge pass over the listed long segments (those of more than
// min_n keys): sorted runs of length `run` pair up into runs of 2*run.
// Each item makes SEG_MERGE_ITEM consecutive outputs of a pair of runs A, B:
// the merge-path splits of its first and last diagonal (the number of A keys
// among the outputs before it) come from a 128-way parallel search each (a
// few rounds of independent loads instead of a chain of dependent ones),
// then the item's slices of A and B (together SEG_MERGE_ITEM keys) are
// loaded into LDS coalesced, merged there four outputs per thread and
// written out contiguously. Keys are distinct.
// The split of diagonal d of A (na keys) and B (nb keys): the first i in
// [max(0, d - nb), min(d, na)) with !(A[i] < B[d - i - 1]), else the upper
// end. 128 threads (q = 0..127) probe evenly spaced i per round, so a range
// of r + 1 candidates takes 1 + ceil(log128(r + 1)) rounds; every thread of
// the work-group runs `rounds` rounds (two searches side by side, slot 0
// and 1, each over two waves), so the barriers match.
ATQ_DEV uint32_t seg_split_search(const uint64_t *A, uint32_t na, const uint64_t *B, uint32_t nb,
                                  uint32_t d, uint32_t q, int rounds, uint32_t *cnt_lds,
                                  int slot) {
    uint32_t lo = d > nb ? d - nb : 0, hi = min(d, na);
    for (int r = 0; r < rounds; r++) {
        const uint32_t step = (hi - lo + 127) / 128;
        const uint32_t i = lo + q * step;
        const bool pred = hi > lo && i < hi && A[i] < B[d - i - 1];
        const uint64_t bal = __builtin_amdgcn_ballot_w64(pred);
        if ((threadIdx.x & 63) == 0) cnt_lds[slot * 2 + ((threadIdx.x >> 6) & 1)] = (uint32_t)__popcll(bal);
        __syncthreads();
        const uint32_t c = cnt_lds[slot * 2] + cnt_lds[slot * 2 + 1];
        __syncthreads();
        if (hi > lo) {
            const uint32_t nlo = c ? lo + (c - 1) * step + 1 : lo;
            hi = min(lo + c * step, hi);
            lo = nlo;
        }
    }
    return lo;
}

template <typename Ext>
__global__ void __launch_bounds__(SEG_MERGE_NT)
k_seg_merge(const Ext *__restrict__ segs, SegLong lng, uint32_t max_items, uint32_t run,
            const uint64_t *__restrict__ kin, uint64_t *__restrict__ kout, uint32_t min_n) {
    static_assert(SEG_MERGE_NT == 256 && SEG_MERGE_ITEM == 4 * SEG_MERGE_NT, "");
    __shared__ uint64_t buf[SEG_MERGE_ITEM];
    __shared__ uint32_t cnt_lds[4];
    __shared__ uint32_t split[2];
    const uint32_t nl = min(*lng.n, lng.cap);
    const uint64_t items = (uint64_t)nl * max_items;
    const uint32_t t = threadIdx.x, half = t >> 7, q = t & 127;
    for (uint64_t it = blockIdx.x; it < items; it += gridDim.x) {  // item-major
        const Ext e = segs[lng.list[it % nl]];
        const uint32_t n = e.count, ob = (uint32_t)(it / nl) * SEG_MERGE_ITEM;
        if (n <= run || ob >= n || n <= min_n) continue;
        const uint64_t *c = kin + e.start;
        const uint32_t base = ob / (2 * run) * (2 * run);
        const uint32_t na = min(run, n - base), nb = n - base > run ? min(run, n - base - run) : 0u;
        const uint64_t *A = c + base, *B = c + base + run;
        const uint32_t d0 = ob - base, d1 = min(d0 + SEG_MERGE_ITEM, na + nb);
        // the two splits side by side, one per half of the work-group
        int rounds = 1;
        for (uint32_t x = run + 1; x > 128; x = (x + 127) / 128) rounds++;
        const uint32_t s = seg_split_search(A, na, B, nb, half ? d1 : d0, q, rounds, cnt_lds, half);
        if (q == 0) split[half] = s;
        __syncthreads();
        const uint32_t i0 = split[0], i1 = split[1], j0 = d0 - i0, j1 = d1 - i1;
        const uint32_t ma = i1 - i0, mb = j1 - j0;
        for (uint32_t k = t; k < ma + mb; k += SEG_MERGE_NT)
            buf[k] = k < ma ? A[i0 + k] : B[j0 + (k - ma)];
        __syncthreads();
        // four outputs per thread: merge path inside the item's slices
        const uint32_t o = 4 * t;
        if (o < ma + mb) {
            uint32_t lo = o > mb ? o - mb : 0, hi = min(o, ma);
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (buf[mid] < buf[ma + o - mid - 1]) lo = mid + 1;
                else hi = mid;
            }
            uint32_t ia = lo, ib = o - lo;
            uint64_t out[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint64_t av = ia < ma ? buf[ia] : SEG_PAD;
                const uint64_t bv = ib < mb ? buf[ma + ib] : SEG_PAD;
                const bool ta = av < bv;
                out[k] = ta ? av : bv;
                ia += ta ? 1u : 0u;
                ib += ta ? 0u : 1u;
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (o + k < ma + mb) kout[e.start + base + d0 + o + k] = out[k];
        }
        __syncthreads();
    }
}

// The chunk sort of every segment: segments of at most SEG_CHUNK keys are
// sorted and finished; longer ones have their chunks sorted into keys0 and
// are listed in lng (n zero on entry).
template <typename KeyGen, typename Finish, typename Ext>
int seg_sort_chunks(hipStream_t s, KeyGen gen, Finish fin, const Ext *segs,
                    const uint32_t *n_segs_dev, uint32_t cap_segs, uint32_t max_seg,
                    uint64_t *keys0, SegLong lng) {
    k_seg_chunks<KeyGen, Finish, Ext><<<1024, SEG_NT, 0, s>>>(gen, fin, segs, n_segs_dev, cap_segs,
                                                               seg_max_chunks(max_seg), keys0, lng);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

// The merges of the listed segments of more than min_n keys after
// seg_sort_chunks, in LDS first when lds (their sorted keys end in
// seg_sorted_keys(n, keys0, keys1, lds)).
template <typename Ext>
int seg_sort_long(hipStream_t s, const Ext *segs, uint32_t max_seg, uint64_t *keys0,
                  uint64_t *keys1, SegLong lng, bool lds, uint32_t min_n = 0) {
    if (lds && max_seg > min_n)
        k_seg_merge_lds<Ext><<<256, SEG_NT, 0, s>>>(
            segs, lng, (max_seg + SEG_LDS_KEYS - 1) / SEG_LDS_KEYS, keys0, keys1, min_n);
    // hop h reads keys0 when h is even (the chunks, or the LDS merge's input)
    const uint32_t max_items = (max_seg + SEG_MERGE_ITEM - 1) / SEG_MERGE_ITEM;
    uint32_t hop = lds ? 1u : 0u;
    for (uint32_t run = lds ? SEG_LDS_KEYS : SEG_CHUNK; run < max_seg; run <<= 1, hop++)
        k_seg_merge<Ext><<<1024, SEG_MERGE_NT, 0, s>>>(segs, lng, max_items, run,
                                                         (hop & 1) ? keys1 : keys0,
                                                         (hop & 1) ? keys0 : keys1, min_n);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

// Sorts every segment; max_seg bounds the segment length (longer segments
// are left partly sorted: callers bound it by construction). lng.n must be
// zero on entry; on return the list names the segments longer than
// SEG_CHUNK.
template <typename KeyGen, typename Finish, typename Ext>
int seg_sort(hipStream_t s, KeyGen gen, Finish fin, const Ext *segs, const uint32_t *n_segs_dev,
             uint32_t cap_segs, uint32_t max_seg, uint64_t *keys0, uint64_t *keys1, SegLong lng,
             bool lds) {
    if (int r = seg_sort_chunks(s, gen, fin, segs, n_segs_dev, cap_segs, max_seg, keys0, lng))
        return r;
    return seg_sort_long(s, segs, max_seg, keys0, keys1, lng, lds);
}

}  // namespace atq
