// records.hpp - device-side semantics of the reference's record types.
#pragma once
#include "common.hpp"

namespace atq {

// include/boundary_point.h:53-54, 63-95
ATQ_DEV uint32_t x_with_dx(uint16_t px) {
    return (uint16_t)(2 * (px & COORD_VALUE_MASK) + half_pixel_dx(px));
}
ATQ_DEV uint32_t y_with_dy(uint16_t px, uint16_t py) {
    return (uint16_t)(2 * (py & COORD_VALUE_MASK) + half_pixel_dy(px));
}
ATQ_DEV bool black_to_white(uint16_t py) { return (py & COLOR_DIRECTION_MASK) != 0; }
ATQ_DEV int grad_x(uint16_t px, uint16_t py) {
    return black_to_white(py) ? half_pixel_dx(px) : -half_pixel_dx(px);
}
ATQ_DEV int grad_y(uint16_t px, uint16_t py) {
    return black_to_white(py) ? half_pixel_dy(px) : -half_pixel_dy(px);
}

// include/cluster_bounds.h:47-49 (float*0.5f, double add, back to float)
ATQ_DEV float bounds_cx(const atq_cluster_bounds &b) {
    return (float)((double)((float)(b.x_min + b.x_max) * 0.5f) + 0.05118);
}
ATQ_DEV float bounds_cy(const atq_cluster_bounds &b) {
    return (float)((double)((float)(b.y_min + b.y_max) * 0.5f) + -0.028581);
}
// include/cluster_bounds.h:51-66
ATQ_DEV float bounds_dot(const atq_cluster_bounds &b) {
    int ints = (b.x_min + b.x_max) * b.gx_sum + (b.y_min + b.y_max) * b.gy_sum;
    float cxcy = (float)((double)ints * 0.5);
    float gxc = (float)(0.05118 * (double)(float)b.gx_sum);
    float gyc = (float)(-0.028581 * (double)(float)b.gy_sum);
    float r = 2.0f * (float)b.pxgx_pygy_sum;
    r = r - cxcy;
    r = r - gxc;
    r = r - gyc;
    return r;
}

// ValidBlobFilter(width, height), include/cluster_bounds.h:85-120
ATQ_DEV bool valid_blob(const atq_cluster_bounds &b, uint32_t w, uint32_t h) {
    const uint64_t max_px = 2ull * (2ull * w + 2ull * h);
    if (b.count < 5 || b.count < 24 || (uint64_t)b.count > max_px) return false;
    int area = (b.x_max - b.x_min) * (b.y_max - b.y_min);
    if ((uint64_t)area < 4) return false;
    return !(bounds_dot(b) < 0.0f);
}

// include/boundary_point.h:48-61
ATQ_DEV float calc_theta(uint16_t px, uint16_t py, float cx, float cy) {
    float dx = (float)x_with_dx(px) - cx;
    float dy = (float)y_with_dy(px, py) - cy;
    bool a = dx > 0, b = dy > 0;
    float q = a ? (b ? 65536.0f : 0.0f) : (b ? 131072.0f : -65536.0f);
    if (a == b) return q + dy / dx;
    return q - dx / dy;
}

// include/line_fit_point.h:50-86, including the row index taken from rx
// (`uint32_t y = rx / 2`).
ATQ_DEV atq_line_fit_point initial_linefit(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                                           const uint8_t *__restrict__ grey) {
    int64_t rx = (int64_t)x_with_dx(px) + 1;
    int64_t ry = (int64_t)y_with_dy(px, py) + 1;
    uint32_t x = (uint32_t)(rx / 2);
    uint32_t y = (uint32_t)(rx / 2);
    int64_t W = 1;
    if (x > 0 && x + 1 < w && y > 0 && y + 1 < h) {
        int gx = (int)grey[(size_t)y * w + x + 1] - (int)grey[(size_t)y * w + x - 1];
        int gy = (int)grey[(size_t)(y + 1) * w + x] - (int)grey[(size_t)(y - 1) * w + x];
        W += (int64_t)hypot_f((float)gx, (float)gy);
    }
    atq_line_fit_point r;
    r.Mx = W * rx;
    r.My = W * ry;
    r.Mxx = W * rx * rx;
    r.Myy = W * ry * ry;
    r.Mxy = W * rx * ry;
    r.W = W;
    r.a = 0;
    r.b = 0;
    return r;
}

// Angle-sort key: slope, then the reference's compaction order
// (plane, y, x), then the colour-direction bit (never needed to break ties:
// (plane, y, x) is unique per record). 63 bits.
ATQ_DEV uint64_t angle_key(uint16_t px, uint16_t py, float slope) {
    uint64_t lo = ((uint64_t)(px >> 14) << 29) | ((uint64_t)(py & COORD_VALUE_MASK) << 15) |
                  ((uint64_t)(px & COORD_VALUE_MASK) << 1) | (uint64_t)(py >> 15);
    return ((uint64_t)float_key(slope) << 32) | lo;
}
ATQ_DEV void angle_key_decode(uint64_t k, atq_cluster_point &p) {
    uint32_t lo = (uint32_t)k;
    uint16_t hp = (lo >> 29) & 3u, y = (lo >> 15) & COORD_VALUE_MASK, x = (lo >> 1) & COORD_VALUE_MASK;
    p.packed_x = (uint16_t)(x | (hp << 14));
    p.packed_y = (uint16_t)(y | ((lo & 1u) << 15));
    p.slope = float_from_key((uint32_t)(k >> 32));
}

}  // namespace atq
