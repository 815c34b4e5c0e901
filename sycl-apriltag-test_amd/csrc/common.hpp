// common.hpp - shared device helpers of the gfx950 quad front-end.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "atq.h"

// All kernels are compiled with -ffp-contract=off: every float expression
// rounds like the reference's IEEE operations, with no fused multiply-add.
#pragma clang fp contract(off)

#define ATQ_DEV __device__ __forceinline__

namespace atq {

constexpr uint32_t LABEL_PIXEL_MASK = 0x80000000u;
constexpr uint32_t LABEL_VALUE_MASK = 0x7FFFFFFFu;
constexpr uint16_t COLOR_DIRECTION_MASK = 1u << 15;
constexpr uint16_t COORD_VALUE_MASK = (1u << 14) - 1;
// (1 << 15) & (1 << 14): the reference's half-pixel mask is 0, so every
// record reads back as RIGHT: dx = 1, dy = 0 (include/boundary_point.h:16).
constexpr uint16_t HALF_PIXEL_MASK = (1u << 15) & (1u << 14);

constexpr int WAVE = 64;

// overflow flag bits in atq_counts::overflow
constexpr uint32_t OVF_BOUNDARY = 1u << 0;
constexpr uint32_t OVF_CORNERS = 1u << 1;
constexpr uint32_t OVF_SCAN = 1u << 2;  // a look-back wait gave up: results void

// The record's half-pixel: always RIGHT (0) as the reference decodes it; in
// fixed mode (fx) both bits, the plane the record came from (0 RIGHT,
// 1 BOTTOM_LEFT, 2 BOTTOM, 3 BOTTOM_RIGHT; include/boundary_point.h:8-28).
ATQ_DEV uint32_t half_pixel(uint16_t packed_x, bool fx) {
    return fx ? ((uint32_t)packed_x >> 14) & 3u : (uint32_t)(packed_x & HALF_PIXEL_MASK) >> 14;
}
ATQ_DEV int half_pixel_dx(uint16_t packed_x, bool fx = false) {
    switch (half_pixel(packed_x, fx)) {
    case 0: return 1;
    case 1: return -1;
    case 2: return 0;
    default: return 1;
    }
}
ATQ_DEV int half_pixel_dy(uint16_t packed_x, bool fx = false) {
    return half_pixel(packed_x, fx) == 0 ? 0 : 1;
}

// glibc's hypotf: the binary64 root of the exactly-rounded binary64 sum.
ATQ_DEV float hypot_f(float a, float b) {
    double da = a, db = b;
    double s = da * da + db * db;
    return (float)__builtin_sqrt(s);
}

ATQ_DEV int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
ATQ_DEV int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
ATQ_DEV int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }

// Wave-level helpers (64 lanes).
ATQ_DEV uint32_t lane_id() { return __lane_id(); }
ATQ_DEV uint64_t lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Exclusive prefix of a small per-lane count within the wave; *total gets
// the wave sum.
ATQ_DEV uint32_t wave_exclusive_sum(uint32_t v, uint32_t *total) {
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t y = __shfl_up(x, o, WAVE);
        if ((int)__lane_id() >= o) x += y;
    }
    *total = __shfl(x, WAVE - 1, WAVE);
    return x - v;
}

// Work-groups are handed to the 8 XCDs (each with its own L2) round-robin in
// launch order. For a grid of 8 * ceil(n / 8) work-groups this maps work-group
// b to tile xcd_tile(b, n), so that each XCD takes one contiguous run of
// tiles and neighbouring tiles share that XCD's L2 (halo rows and columns are
// then fetched from HBM once). Returns n (no tile) for the padding groups.
constexpr uint32_t N_XCD = 8;
ATQ_DEV uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t per = (n + N_XCD - 1) / N_XCD;
    const uint32_t t = (b % N_XCD) * per + b / N_XCD;
    return t < n ? t : n;
}
__host__ __device__ inline uint32_t xcd_grid(uint32_t n) { return (n + N_XCD - 1) / N_XCD * N_XCD; }

// Orderable image of a float (no NaN is ever sorted here): ascending u32
// order equals ascending float order.
ATQ_DEV uint32_t float_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
ATQ_DEV float float_from_key(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

}  // namespace atq

#define ATQ_HIP_CHECK(expr)                                   \
    do {                                                      \
        hipError_t e_ = (expr);                               \
        if (e_ != hipSuccess) return ATQ_ERR_HIP;             \
    } while (0)

static inline unsigned atq_div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
