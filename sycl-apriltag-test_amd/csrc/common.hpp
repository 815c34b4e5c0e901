// common.hpp - shared device helpers of the gfx950 quad front-end.
#pragma once
#include <type_traits>
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

// DPP lane moves (gfx9-family encodings), every lane of the wave active:
// row_shr:n (0x110 + n) moves a value n lanes up within each row of 16
// lanes (lanes with no source in the row read 0); row_bcast:15 (0x142, rows
// 1 and 3) hands lane 15 of each row to the next row, row_bcast:31 (0x143,
// rows 2 and 3) hands lane 31 to the upper half; the rows outside the mask
// read 0. One VALU move each, against a round trip through the LDS
// crossbar for a lane shuffle.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
ATQ_DEV uint32_t dpp_move(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, 0xF, BOUND_ZERO);
}
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
ATQ_DEV uint64_t dpp_move(uint64_t x) {
    const uint32_t lo = dpp_move<CTRL, ROW_MASK, BOUND_ZERO>((uint32_t)x);
    const uint32_t hi = dpp_move<CTRL, ROW_MASK, BOUND_ZERO>((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The value of the lane one below (wave_shr:1, 0x138) or one above
// (wave_shl:1, 0x130) across the whole wave, crossing the rows; lane 0
// (below) and lane 63 (above) read 0. Every lane of the wave active.
ATQ_DEV uint32_t lane_below(uint32_t x) { return dpp_move<0x138, 0xF, true>(x); }
ATQ_DEV uint32_t lane_above(uint32_t x) { return dpp_move<0x130, 0xF, true>(x); }

// Inclusive sum over the 64 lanes of a wave in lane order (wrapping
// arithmetic; 32- or 64-bit integers), every lane active.
template <typename T>
ATQ_DEV T wave_inclusive_sum(T x) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
    using U = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    U u = (U)x;
    u += dpp_move<0x111, 0xF, true>(u);
    u += dpp_move<0x112, 0xF, true>(u);
    u += dpp_move<0x114, 0xF, true>(u);
    u += dpp_move<0x118, 0xF, true>(u);
    u += dpp_move<0x142, 0xA, false>(u);
    u += dpp_move<0x143, 0xC, false>(u);
    return (T)u;
}
// Maximum over the 64 lanes of a wave (unsigned), in lane 63 (the lanes
// below hold partial maxima); every lane active.
ATQ_DEV uint32_t wave_max_to_last(uint32_t u) {
    u = max(u, dpp_move<0x111, 0xF, true>(u));
    u = max(u, dpp_move<0x112, 0xF, true>(u));
    u = max(u, dpp_move<0x114, 0xF, true>(u));
    u = max(u, dpp_move<0x118, 0xF, true>(u));
    u = max(u, dpp_move<0x142, 0xA, false>(u));
    u = max(u, dpp_move<0x143, 0xC, false>(u));
    return u;
}
// The same over the first 16 lanes only (lanes 16-63 get their row's sums).
template <typename T>
ATQ_DEV T row_inclusive_sum(T x) {
    using U = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type;
    U u = (U)x;
    u += dpp_move<0x111, 0xF, true>(u);
    u += dpp_move<0x112, 0xF, true>(u);
    u += dpp_move<0x114, 0xF, true>(u);
    u += dpp_move<0x118, 0xF, true>(u);
    return (T)u;
}
// Lane l's value, as a wave-uniform (scalar) value.
ATQ_DEV uint32_t lane_value(uint32_t x, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}
ATQ_DEV uint64_t lane_value(uint64_t x, int l) {
    return ((uint64_t)lane_value((uint32_t)(x >> 32), l) << 32) | lane_value((uint32_t)x, l);
}
ATQ_DEV int64_t lane_value(int64_t x, int l) { return (int64_t)lane_value((uint64_t)x, l); }

// One step of wave_min_u64: the minimum of x and the DPP-moved x (all ones
// where the move finds no source lane).
template <int CTRL, int ROW_MASK>
ATQ_DEV uint64_t dpp_min_step(uint64_t x) {
    const uint32_t lo =
        (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)x, CTRL, ROW_MASK, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)(x >> 32), CTRL,
                                                              ROW_MASK, 0xF, false);
    const uint64_t y = ((uint64_t)hi << 32) | lo;
    return y < x ? y : x;
}
// Minimum over the 64 lanes of a wave (unsigned 64-bit), in every lane;
// every lane active.
ATQ_DEV uint64_t wave_min_u64(uint64_t v) {
    v = dpp_min_step<0x111, 0xF>(v);
    v = dpp_min_step<0x112, 0xF>(v);
    v = dpp_min_step<0x114, 0xF>(v);
    v = dpp_min_step<0x118, 0xF>(v);
    v = dpp_min_step<0x142, 0xA>(v);
    v = dpp_min_step<0x143, 0xC>(v);
    return lane_value(v, WAVE - 1);
}

// Exclusive prefix of a small per-lane count within the wave; *total gets
// the wave sum. Every lane active.
ATQ_DEV uint32_t wave_exclusive_sum(uint32_t v, uint32_t *total) {
    const uint32_t x = wave_inclusive_sum(v);
    *total = lane_value(x, WAVE - 1);
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
