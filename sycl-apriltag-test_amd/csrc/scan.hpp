// scan.hpp - work-group and device-wide scans over device-resident counts.
//
// Device-wide scans run in three kernels over a fixed grid of G work-groups
// (the element count is read from device memory, so nothing waits on the
// host): per-chunk totals, an exclusive scan of the G totals, then each
// work-group rescans its chunk in rounds of 256 and hands every element its
// exclusive prefix. Chunks are contiguous, so the order is the element order.
#pragma once
#include "common.hpp"

namespace atq {

constexpr int SCAN_THREADS = 256;
constexpr unsigned SCAN_GROUPS = 1024;

// Exclusive scan over the NT threads of a work-group (lds holds NT/64
// entries); *total = group sum.
template <typename T, int NT = 256>
ATQ_DEV T block_exclusive_scan(T v, T *lds, T *total) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const T x = wave_inclusive_sum(v);
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T before = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        T t = lds[w];
        if (w < wave) before += t;
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return before + x - v;
}

template <typename T>
ATQ_DEV T block_reduce_sum(T v, T *lds4) {
    T total;
    block_exclusive_scan<T>(v, lds4, &total);
    return total;
}

ATQ_DEV void chunk_range(uint32_t n, uint32_t g, uint32_t ngroups, uint32_t *b, uint32_t *e) {
    uint32_t per = (n + ngroups - 1) / ngroups;
    per = (per + SCAN_THREADS - 1) / SCAN_THREADS * SCAN_THREADS;
    uint64_t bb = (uint64_t)g * per, ee = bb + per;
    *b = (uint32_t)(bb < n ? bb : n);
    *e = (uint32_t)(ee < n ? ee : n);
}

// F: uint32_t operator()(uint32_t i) const   (value of element i)
template <typename F>
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_partials(const uint32_t *count, uint32_t count_cap, F f, uint32_t *partials) {
    __shared__ uint32_t lds4[4];
    uint32_t n = count ? min(*count, count_cap) : count_cap;
    uint32_t b, e;
    chunk_range(n, blockIdx.x, gridDim.x, &b, &e);
    uint32_t acc = 0;
    for (uint32_t i = b + threadIdx.x; i < e; i += SCAN_THREADS) acc += f(i);
    uint32_t t = block_reduce_sum<uint32_t>(acc, lds4);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// One work-group of 1024: exclusive scan of `ng` (<= 1024) partials in
// place; the total goes to *total and to partials[ng].
template <int D = 0>
__global__ void __launch_bounds__(1024) k_scan_top(uint32_t *partials, uint32_t ng, uint32_t *total) {
    __shared__ uint32_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    uint32_t v = t < (int)ng ? partials[t] : 0u;
    const uint32_t x = wave_inclusive_sum(v);
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, sum = 0;
    for (int w = 0; w < 16; w++) {
        if (w < wave) before += wsum[w];
        sum += wsum[w];
    }
    if (t < (int)ng) partials[t] = before + x - v;
    if (t == 0) {
        partials[ng] = sum;
        if (total) *total = sum;
    }
}

// O: void operator()(uint32_t i, uint32_t exclusive_prefix, uint32_t value) const
template <typename F, typename O>
__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_apply(const uint32_t *count, uint32_t count_cap, F f, const uint32_t *partials, O o) {
    __shared__ uint32_t lds4[4];
    uint32_t n = count ? min(*count, count_cap) : count_cap;
    uint32_t b, e;
    chunk_range(n, blockIdx.x, gridDim.x, &b, &e);
    uint32_t carry = partials[blockIdx.x];
    for (uint32_t base = b; base < e; base += SCAN_THREADS) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < e ? f(i) : 0u;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan<uint32_t>(v, lds4, &tot);
        if (i < e) o(i, carry + ex, v);
        carry += tot;
    }
}

// Single-pass form (decoupled look-back): work-group p scans its tile of
// LB_TILE elements, publishes the tile's sum, then walks back over the
// flags of the tiles before it (64 at a time, one per lane) until it meets
// an inclusive prefix, and publishes its own. Flag words: status in bits
// 62-63 (1: tile sum, 2: inclusive prefix, 0: not yet), value in bits 0-31.
// Work-groups are dispatched in index order, so every tile waited on is
// running or done; the wait is still bounded (LB_SPIN_LIMIT reads), after
// which *error gets OVF_SCAN and the scan's results are void.
// Each call uses one of two flag arrays and clears the other for the next
// call (both must be zero before the first call of a sequence).
constexpr int LB_THREADS = 256;
constexpr int LB_ITEMS = 8;
constexpr uint32_t LB_TILE = LB_THREADS * LB_ITEMS;
// Scans of at most LB_SMALL_CAP elements use tiles of half the size: twice
// the work-groups, each with half the work, shortens the chain of the
// 1080p frame's scans (14.3 against 16.6 µs per frame, measured); at 12 MP
// the larger tiles win, the launch retiring half as many idle work-groups.
constexpr uint32_t LB_SMALL_CAP = 1u << 24;
constexpr uint64_t LB_SUM = 1ull << 62, LB_PREFIX = 2ull << 62;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 22;

ATQ_DEV uint64_t lb_load(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
ATQ_DEV void lb_store(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by the 64 lanes of one wave of work-group p: publishes p's sum,
// walks back to an inclusive prefix, publishes p's own and returns the sum
// of all tiles before p (every lane gets it).
ATQ_DEV uint32_t lb_lookback(uint64_t *flags, uint32_t p, uint32_t tile_sum, uint32_t *error) {
    const uint32_t lane = threadIdx.x & 63;
    if (p == 0) {
        if (lane == 0) lb_store(&flags[0], LB_PREFIX | tile_sum);
        return 0;
    }
    if (lane == 0) lb_store(&flags[p], LB_SUM | tile_sum);
    uint32_t excl = 0, spins = 0;
    int64_t q = (int64_t)p - 1;
    while (true) {
        const int64_t idx = q - (int64_t)lane;
        const uint64_t fl = idx >= 0 ? lb_load(&flags[idx]) : LB_PREFIX;
        const uint64_t st = fl >> 62;
        const uint64_t pre = __ballot(st == 2), zero = __ballot(st == 0);
        const int first_pre = pre ? __ffsll((long long)pre) - 1 : 64;
        const uint64_t before = first_pre == 64 ? ~0ull : ((1ull << first_pre) - 1);
        if ((zero & before) && ++spins < LB_SPIN_LIMIT) continue;
        const uint32_t c = (int)lane <= first_pre ? (uint32_t)fl : 0u;
        excl += lane_value(wave_inclusive_sum(c), 63);
        if (first_pre < 64 || spins >= LB_SPIN_LIMIT) break;
        q -= 64;
    }
    if (spins >= LB_SPIN_LIMIT && lane == 0 && error) atomicOr(error, OVF_SCAN);
    if (lane == 0) lb_store(&flags[p], LB_PREFIX | (uint32_t)(excl + tile_sum));
    return excl;
}

template <typename F, typename O, int ITEMS>
__global__ void __launch_bounds__(LB_THREADS)
k_scan_lookback(const uint32_t *count, uint32_t count_cap, F f, O o, uint64_t *flags,
                uint64_t *next_flags, uint32_t n_flags, uint32_t *total, uint32_t *error) {
    __shared__ uint32_t lds16[LB_THREADS / 64];
    __shared__ uint32_t s_prefix;
    for (uint32_t i = blockIdx.x * LB_THREADS + threadIdx.x; i < n_flags;
         i += gridDim.x * LB_THREADS)
        next_flags[i] = 0;
    const uint32_t n = count ? min(*count, count_cap) : count_cap;
    constexpr uint32_t TILE = LB_THREADS * ITEMS;
    const uint32_t p = blockIdx.x, b = p * TILE;
    if (b >= n) {
        if (p == 0 && threadIdx.x == 0) *total = 0;
        return;
    }
    const uint32_t i0 = b + threadIdx.x * ITEMS;
    uint32_t v[ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        v[k] = i0 + k < n ? f(i0 + k) : 0u;
        sum += v[k];
    }
    uint32_t tile_sum;
    const uint32_t ex = block_exclusive_scan<uint32_t, LB_THREADS>(sum, lds16, &tile_sum);
    if (threadIdx.x < 64) {
        const uint32_t excl = lb_lookback(flags, p, tile_sum, error);
        if (threadIdx.x == 0) s_prefix = excl;
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
    if (threadIdx.x == 0 && b + TILE >= n) *total = prefix + tile_sum;
    uint32_t run = prefix + ex;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        if (i0 + k < n) o(i0 + k, run, v[k]);
        run += v[k];
    }
}

// Where a device scan keeps its state. flags null: the three-kernel form
// (scratch: SCAN_GROUPS + 1 u32). Otherwise the look-back form: flags holds
// two arrays of flag_count u64 (zero before the first scan), `calls`
// counts the scans launched so far (it picks the array), error receives
// the look-back failure bit.
struct ScanCtx {
    uint32_t *scratch = nullptr;
    uint64_t *flags = nullptr;
    uint32_t flag_count = 0;
    uint32_t calls = 0;
    uint32_t *error = nullptr;
};

// Scans min(*count, count_cap) elements (count == nullptr: exactly
// count_cap): o(i, exclusive prefix, value) for every element; *total
// receives the sum (device memory).
// ONE_ITEM: one element per thread in the look-back form, for flag
// functors that do a whole item's work (the elements' latencies then
// overlap across the threads instead of adding up in each).
template <typename F, typename O, bool ONE_ITEM = false>
int device_scan(hipStream_t s, const uint32_t *count, uint32_t count_cap, F f, O o, ScanCtx &sc,
                uint32_t *total) {
    if (sc.flags) {
        const bool small = count_cap <= LB_SMALL_CAP;
        const uint32_t tiles =
            count_cap / (ONE_ITEM ? LB_THREADS : small ? LB_TILE / 2 : LB_TILE) + 1;
        if (tiles > sc.flag_count) return ATQ_ERR_CAPACITY;
        uint64_t *cur = sc.flags + (size_t)(sc.calls & 1) * sc.flag_count;
        uint64_t *next = sc.flags + (size_t)((sc.calls + 1) & 1) * sc.flag_count;
        sc.calls++;
        if (ONE_ITEM)
            k_scan_lookback<F, O, 1><<<tiles, LB_THREADS, 0, s>>>(
                count, count_cap, f, o, cur, next, sc.flag_count, total, sc.error);
        else if (small)
            k_scan_lookback<F, O, LB_ITEMS / 2><<<tiles, LB_THREADS, 0, s>>>(
                count, count_cap, f, o, cur, next, sc.flag_count, total, sc.error);
        else
            k_scan_lookback<F, O, LB_ITEMS><<<tiles, LB_THREADS, 0, s>>>(
                count, count_cap, f, o, cur, next, sc.flag_count, total, sc.error);
        return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
    }
    k_scan_partials<F><<<SCAN_GROUPS, SCAN_THREADS, 0, s>>>(count, count_cap, f, sc.scratch);
    k_scan_top<0><<<1, 1024, 0, s>>>(sc.scratch, SCAN_GROUPS, total);
    k_scan_apply<F, O><<<SCAN_GROUPS, SCAN_THREADS, 0, s>>>(count, count_cap, f, sc.scratch, o);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
