// sort.hpp - radix sort launcher and the in-work-group bitonic sort.
#pragma once
#include "scan.hpp"

namespace atq {

constexpr unsigned RS_GROUPS = 1024;     // most chunks per pass
constexpr unsigned RS_MAX_DIGITS = 2048; // 11-bit digits

// How a frame of npx pixels sorts its packed pair keys.
struct RadixPlan {
    int label_bits;   // lb: labels are below 2^lb
    int passes;       // over the 2*lb (+ extra) packed key bits
    int digit_bits;   // 8, 10 or 11
};
// Keys of 2*lb + extra_bits bits (labels below npx).
RadixPlan radix_plan(size_t npx, int extra_bits);

// Sorts (keys, vals) by the low passes*digit_bits key bits, stably. `hist`
// holds RS_MAX_DIGITS * RS_GROUPS u32, `totals` RS_MAX_DIGITS u32. On
// return the sorted data are in the alt buffers when *result_in_alt is true.
int radix_sort_packed(hipStream_t s, const RadixPlan &p, uint64_t *keys, uint32_t *vals,
                      uint64_t *keys_alt, uint32_t *vals_alt, const uint32_t *count, uint32_t cap,
                      uint32_t *hist, uint32_t *totals, bool *result_in_alt);

// The detector's sort of its boundary records (sort.hip): keys built from
// r, the rank of the record's cluster pair among the frame's valid pairs,
// either r << 2 | plane with the records in raster order (rank_shift 2), or
// r alone with the records arriving plane by plane, each plane in raster
// order (rank_shift 0; one bit-pair shorter, used when that saves a pass);
// either way a stable sort gives the reference's (pair, plane, y, x) order.
// Sorted with RK_BITS-bit digits in *passes (1..RK_MAX_PASSES) passes from
// X or Y (see rank_sort) into Y. Every kept record belongs to a valid pair,
// so cluster r is rank r: the last pass writes the packed key
// (xlat[r] << 2 | plane) and also each record as the cluster point of
// cluster r (slope about centre[r], which the ranking wrote with the bounds
// and extents). *passes == 0 selects the fallback: packed keys (with the
// plane) in X, sorted into Y by one persistent kernel that then groups the
// records into segments, reduces and filters them and assigns the cluster
// points itself.
constexpr int RK_BITS = 11;
constexpr uint32_t RK_MAX_PASSES = 2;
constexpr unsigned RS_FULL_GROUPS = 128;  // work-groups of the fallback
// Where the sort leaves the clusters (src/main.cpp:571-675), and the
// fallback's segment tables (cap entries each).
struct RsClusters {
    atq_cluster_bounds *bounds;
    atq_cluster_extents *ext;
    float2 *centre;           // rank mode: each cluster's (cx, cy) from its bounds
    atq_cluster_point *cpts;  // records in cluster order, with their slopes
    uint32_t *cidx;           // fallback: cluster of each
    uint32_t *seg_id, *seg_start, *valid, *seg_cluster;
    atq_cluster_bounds *seg_bounds;
    uint32_t *part;  // 3 * RS_FULL_GROUPS
    atq_counts *counts;
    uint32_t w, h, fx;
};
struct RsDyn {
    uint64_t *kx;
    uint32_t *vx;
    uint64_t *ky;
    uint32_t *vy;
    const uint32_t *passes;  // device: rank passes, 0 = full-key fallback
    const uint32_t *shift;   // device: rank_shift (rank = key >> shift)
    const uint64_t *xlat;    // rank -> packed pair key
    RsClusters cl;
};
// Radix passes over the rank keys of v valid pairs: ranks below v take
// bit_length(v - 1) bits, plus `extra` (the plane's 2 bits in raster mode);
// 0 when that needs more than RK_MAX_PASSES digits.
__host__ __device__ inline uint32_t rank_passes(uint32_t v, uint32_t extra) {
    uint32_t bits = extra;
    for (uint32_t r = v > 0 ? v - 1 : 0; r; r >>= 1) bits++;
    const uint32_t np = bits ? (bits + RK_BITS - 1) / RK_BITS : 1u;
    return np <= RK_MAX_PASSES ? np : 0u;
}
// Raster mode (keys r << 2 | plane, records in raster order) when it needs
// no more passes than plane-major mode (keys r, records plane by plane):
// the raster layout is the cheaper one to produce.
__host__ __device__ inline uint32_t rank_shift_for(uint32_t v) {
    return rank_passes(v, 2) == rank_passes(v, 0) ? 2u : 0u;
}
// Launches RK_MAX_PASSES rank passes (idle beyond *passes), the last of
// which also holds the fallback over full_bits-bit packed keys (idle unless
// *passes == 0). `bar` is a zeroed word (the fallback's grid barrier);
// `error` gets OVF_SCAN if a barrier wait gives up.
int rank_sort(hipStream_t s, uint32_t full_bits, const RsDyn &d, const uint32_t *count,
              uint32_t cap, uint32_t *hist, uint32_t *totals, uint32_t *bar, uint32_t *error);

// min<<32|max <-> min<<lb|max over count keys (out may alias in).
int launch_pack_keys(hipStream_t s, const uint64_t *in, uint64_t *out, const uint32_t *count,
                     uint32_t cap, uint32_t lb, bool unpack);

// Ascending sort of a[0..n) by a work-group of `nthreads` threads with the
// bitonic network in its "all comparators ascending" form, so that the
// missing elements a[n..P) of the power-of-two size P act as +infinity and
// are never touched. `a` is LDS or global memory owned by this work-group.
// All threads of the work-group must call it.
template <typename T>
ATQ_DEV void block_bitonic_sort(T *a, uint32_t n, int nthreads, bool global_mem) {
    if (n < 2) return;
    uint32_t P = 1;
    while (P < n) P <<= 1;
    const uint32_t t = threadIdx.x;
    for (uint32_t lk = 1; (1u << lk) <= P; lk++) {
        const uint32_t k = 1u << lk;
        // flip: i in the lower half of each k-block against its mirror
        for (uint32_t c = t; c < P / 2; c += nthreads) {
            uint32_t i = ((c >> (lk - 1)) << lk) | (c & ((k >> 1) - 1));
            uint32_t j = i ^ (k - 1);
            if (j < n) {
                T x = a[i], y = a[j];
                if (y < x) { a[i] = y; a[j] = x; }
            }
        }
        if (global_mem) __threadfence_block();
        __syncthreads();
        for (int ls = (int)lk - 2; ls >= 0; ls--) {
            const uint32_t st = 1u << ls;
            for (uint32_t c = t; c < P / 2; c += nthreads) {
                uint32_t i = ((c >> ls) << (ls + 1)) | (c & (st - 1));
                uint32_t j = i + st;
                if (j < n) {
                    T x = a[i], y = a[j];
                    if (y < x) { a[i] = y; a[j] = x; }
                }
            }
            if (global_mem) __threadfence_block();
            __syncthreads();
        }
    }
}

}  // namespace atq
