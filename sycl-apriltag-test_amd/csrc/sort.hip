// sort.hip - LSD radix sort of (u64 key, u32 value) records (row A9).
//
// Groups the boundary records by cluster-pair key (the reference's
// sort_by_key, src/main.cpp:515-518). Only the key bits that can be set are
// sorted: a pair key is min_label << 32 | max_label with labels below W*H,
// so the passes cover bits [0, lb) and [32, 32 + lb), lb = bit length of
// W*H - 1. Each 8-bit pass is three kernels over a fixed grid of G
// work-groups owning contiguous chunks (the record count is read on the
// device): per-chunk digit histograms, one scan of the digit-major
// histogram table, and a stable scatter that ranks equal digits inside each
// wave with ballots (a 64-lane multisplit) and across the 4 waves in LDS.
// The sort is stable, so equal keys keep their input order.
#include "sort.hpp"

namespace atq {

namespace {
constexpr int RS_THREADS = 256;
constexpr int RS_BUCKETS = 256;
}  // namespace

__global__ void __launch_bounds__(RS_THREADS)
k_rs_hist(const uint64_t *__restrict__ keys, const uint32_t *count, uint32_t cap, uint32_t shift,
          uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[RS_BUCKETS];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint32_t n = min(*count, cap);
    uint32_t b, e;
    chunk_range(n, blockIdx.x, gridDim.x, &b, &e);
    for (uint32_t i = b + threadIdx.x; i < e; i += RS_THREADS)
        atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 0xFFu], 1u);
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the 256*G digit-major histogram table by one work-group.
__global__ void __launch_bounds__(1024) k_rs_scan(uint32_t *hist, uint32_t total_entries) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry_s;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) carry_s = 0;
    __syncthreads();
    for (uint32_t base = 0; base < total_entries; base += 1024) {
        uint32_t i = base + t;
        uint32_t v = i < total_entries ? hist[i] : 0u;
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
        uint32_t carry = carry_s;
        if (i < total_entries) hist[i] = carry + before + x - v;
        __syncthreads();
        if (t == 0) carry_s = carry + sum;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(RS_THREADS)
k_rs_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
             uint64_t *__restrict__ kout, uint32_t *__restrict__ vout, const uint32_t *count,
             uint32_t cap, uint32_t shift, const uint32_t *__restrict__ hist) {
    __shared__ uint32_t wc[4][RS_BUCKETS];   // per-wave digit counts -> prefixes
    __shared__ uint32_t base_off[RS_BUCKETS];
    const int t = threadIdx.x, wave = t >> 6;
    base_off[t] = hist[t * gridDim.x + blockIdx.x];
    uint32_t n = min(*count, cap);
    uint32_t b, e;
    chunk_range(n, blockIdx.x, gridDim.x, &b, &e);
    const uint64_t lt = lanemask_lt();
    for (uint32_t r = b; r < e; r += RS_THREADS) {
#pragma unroll
        for (int w = 0; w < 4; w++) wc[w][t] = 0;
        __syncthreads();
        uint32_t i = r + t;
        bool valid = i < e;
        uint64_t k = valid ? kin[i] : 0ull;
        uint32_t v = valid ? vin[i] : 0u;
        uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            uint64_t bb = __ballot((d >> bit) & 1u);
            m &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        {
            uint32_t run = base_off[t];
#pragma unroll
            for (int w = 0; w < 4; w++) {
                uint32_t c = wc[w][t];
                wc[w][t] = run;
                run += c;
            }
            base_off[t] = run;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = wc[wave][d] + rank;
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
    }
}

int radix_sort_pairs(hipStream_t s, uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                     uint32_t *vals_alt, const uint32_t *count, uint32_t cap,
                     const uint32_t *shifts, int n_shifts, uint32_t *hist, bool *result_in_alt) {
    const unsigned G = RS_GROUPS;
    uint64_t *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    for (int p = 0; p < n_shifts; p++) {
        k_rs_hist<<<G, RS_THREADS, 0, s>>>(ki, count, cap, shifts[p], hist);
        k_rs_scan<<<1, 1024, 0, s>>>(hist, RS_BUCKETS * G);
        k_rs_scatter<<<G, RS_THREADS, 0, s>>>(ki, vi, ko, vo, count, cap, shifts[p], hist);
        uint64_t *tk = ki; ki = ko; ko = tk;
        uint32_t *tv = vi; vi = vo; vo = tv;
    }
    *result_in_alt = (n_shifts & 1) != 0;
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int pair_key_shifts(size_t npx, uint32_t *shifts) {
    int lb = 1;
    while (lb < 32 && ((size_t)1 << lb) < npx) lb++;
    int n = 0;
    for (int sh = 0; sh < lb; sh += 8) shifts[n++] = (uint32_t)sh;
    for (int sh = 0; sh < lb; sh += 8) shifts[n++] = (uint32_t)(32 + sh);
    return n;
}

}  // namespace atq
