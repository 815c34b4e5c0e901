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
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
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
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
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

// Launches the three scan kernels over min(*count, count_cap) elements
// (count == nullptr: exactly count_cap). `scratch` holds SCAN_GROUPS + 1
// u32; *total receives the sum (device memory).
template <typename F, typename O>
int device_scan(hipStream_t s, const uint32_t *count, uint32_t count_cap, F f, O o,
                uint32_t *scratch, uint32_t *total) {
    k_scan_partials<F><<<SCAN_GROUPS, SCAN_THREADS, 0, s>>>(count, count_cap, f, scratch);
    k_scan_top<0><<<1, 1024, 0, s>>>(scratch, SCAN_GROUPS, total);
    k_scan_apply<F, O><<<SCAN_GROUPS, SCAN_THREADS, 0, s>>>(count, count_cap, f, scratch, o);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
