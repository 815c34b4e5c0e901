// records.hpp - device-side semantics of the reference's record types.
#pragma once
#include "common.hpp"

// The record types are layout-identical to the reference's
// (src/main.cpp:435 bit-casts BoundaryPoint to u32, :575 ClusterBounds to
// vec<int64,4>, :794 LineFitPoint to vec<double,8>; include/fit_quad.h).
static_assert(sizeof(atq_extrema) == 2, "vec<u8,2>");
static_assert(sizeof(atq_boundary_point) == 4, "BoundaryPoint");
static_assert(sizeof(atq_cluster_point) == 8, "ClusterPoint");
static_assert(sizeof(atq_cluster_extents) == 8, "ClusterExtents");
static_assert(sizeof(atq_cluster_bounds) == 32, "ClusterBounds");
static_assert(sizeof(atq_line_fit_point) == 64, "LineFitPoint");
static_assert(sizeof(atq_peak_extents) == 8, "PeakExtents");
static_assert(sizeof(atq_corner) == 12, "Corner");
static_assert(sizeof(atq_fitted_quad) == 272, "FittedQuad");
static_assert(sizeof(atq_quad_corners) == 40, "QuadCorners");
static_assert(sizeof(atq_shapes) == 12, "atq_shapes");

namespace atq {

// include/boundary_point.h:53-54, 63-95
// (fx: fixed mode, both half-pixel bits decoded)
ATQ_DEV uint32_t x_with_dx(uint16_t px, bool fx = false) {
    return (uint16_t)(2 * (px & COORD_VALUE_MASK) + half_pixel_dx(px, fx));
}
ATQ_DEV uint32_t y_with_dy(uint16_t px, uint16_t py, bool fx = false) {
    return (uint16_t)(2 * (py & COORD_VALUE_MASK) + half_pixel_dy(px, fx));
}
ATQ_DEV bool black_to_white(uint16_t py) { return (py & COLOR_DIRECTION_MASK) != 0; }
ATQ_DEV int grad_x(uint16_t px, uint16_t py, bool fx = false) {
    return black_to_white(py) ? half_pixel_dx(px, fx) : -half_pixel_dx(px, fx);
}
ATQ_DEV int grad_y(uint16_t px, uint16_t py, bool fx = false) {
    return black_to_white(py) ? half_pixel_dy(px, fx) : -half_pixel_dy(px, fx);
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

// A ClusterBounds reduction in progress (include/cluster_bounds.h:14-83):
// doubled-coordinate bounding box, gradient sums and the moment sum.
struct BoundsAcc {
    uint32_t xmin, xmax, ymin, ymax;
    int32_t gxs, gys;
    int64_t pg;
    ATQ_DEV void init() {
        xmin = ymin = 0xFFFF;
        xmax = ymax = 0;
        gxs = gys = 0;
        pg = 0;
    }
    ATQ_DEV void add(atq_boundary_point p, bool fx) {
        const uint32_t xx = x_with_dx(p.packed_x, fx), yy = y_with_dy(p.packed_x, p.packed_y, fx);
        const int gx = grad_x(p.packed_x, p.packed_y, fx), gy = grad_y(p.packed_x, p.packed_y, fx);
        xmin = min(xmin, xx);
        xmax = max(xmax, xx);
        ymin = min(ymin, yy);
        ymax = max(ymax, yy);
        gxs += gx;
        gys += gy;
        pg += (int64_t)xx * gx + (int64_t)yy * gy;
    }
    ATQ_DEV void merge(const BoundsAcc &o) {
        xmin = min(xmin, o.xmin);
        xmax = max(xmax, o.xmax);
        ymin = min(ymin, o.ymin);
        ymax = max(ymax, o.ymax);
        gxs += o.gxs;
        gys += o.gys;
        pg += o.pg;
    }
    ATQ_DEV void wave_reduce() {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            BoundsAcc t;
            t.xmin = __shfl_xor(xmin, o, 64);
            t.xmax = __shfl_xor(xmax, o, 64);
            t.ymin = __shfl_xor(ymin, o, 64);
            t.ymax = __shfl_xor(ymax, o, 64);
            t.gxs = __shfl_xor(gxs, o, 64);
            t.gys = __shfl_xor(gys, o, 64);
            t.pg = __shfl_xor(pg, o, 64);
            merge(t);
        }
    }
};

// include/boundary_point.h:48-61
ATQ_DEV float calc_theta(uint16_t px, uint16_t py, float cx, float cy, bool fx = false) {
    float dx = (float)x_with_dx(px, fx) - cx;
    float dy = (float)y_with_dy(px, py, fx) - cy;
    bool a = dx > 0, b = dy > 0;
    float q = a ? (b ? 65536.0f : 0.0f) : (b ? 131072.0f : -65536.0f);
    if (a == b) return q + dy / dx;
    return q - dx / dy;
}

// include/line_fit_point.h:50-86, including the row index taken from rx
// (`uint32_t y = rx / 2`; fixed mode: ry / 2): the weight W and the doubled
// coordinates + 1 that every moment of the point is built from.
ATQ_DEV void linefit_terms(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                           const uint8_t *__restrict__ grey, uint32_t *W_out, uint32_t *rx_out,
                           uint32_t *ry_out, bool fx = false) {
    const uint32_t rx = x_with_dx(px, fx) + 1;
    const uint32_t ry = y_with_dy(px, py, fx) + 1;
    const uint32_t x = rx / 2;
    const uint32_t y = (fx ? ry : rx) / 2;
    uint32_t W = 1;
    if (x > 0 && x + 1 < w && y > 0 && y + 1 < h) {
        int gx = (int)grey[(size_t)y * w + x + 1] - (int)grey[(size_t)y * w + x - 1];
        int gy = (int)grey[(size_t)(y + 1) * w + x] - (int)grey[(size_t)(y - 1) * w + x];
        W += (uint32_t)hypot_f((float)gx, (float)gy);
    }
    *W_out = W;
    *rx_out = rx;
    *ry_out = ry;
}

// linefit_terms in two steps: the four grey pixels it may read (0 where it
// would not read), then the terms from them.
ATQ_DEV void linefit_gather(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                            const uint8_t *__restrict__ grey, uint32_t (&g)[4], bool fx = false) {
    const uint32_t rx = x_with_dx(px, fx) + 1;
    const uint32_t x = rx / 2, y = (fx ? y_with_dy(px, py, fx) + 1 : rx) / 2;
    const bool in = x > 0 && x + 1 < w && y > 0 && y + 1 < h;
    const size_t c = in ? (size_t)y * w + x : (size_t)w + 1;  // any in-image pixel when unused
    g[0] = in ? grey[c + 1] : 0u;
    g[1] = in ? grey[c - 1] : 0u;
    g[2] = in ? grey[c + w] : 0u;
    g[3] = in ? grey[c - w] : 0u;
}
ATQ_DEV void linefit_terms_from(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                                const uint32_t (&g)[4], uint32_t *W_out, uint32_t *rx_out,
                                uint32_t *ry_out, bool fx = false) {
    const uint32_t rx = x_with_dx(px, fx) + 1;
    const uint32_t ry = y_with_dy(px, py, fx) + 1;
    const uint32_t x = rx / 2;
    const uint32_t y = (fx ? ry : rx) / 2;
    uint32_t W = 1;
    if (x > 0 && x + 1 < w && y > 0 && y + 1 < h) {
        const int gx = (int)g[0] - (int)g[1];
        const int gy = (int)g[2] - (int)g[3];
        W += (uint32_t)hypot_f((float)gx, (float)gy);
    }
    *W_out = W;
    *rx_out = rx;
    *ry_out = ry;
}

// W <= 1 + hypot(255, 255) = 361 and rx, ry <= 2 * 8192 + 2, so W*rx and
// W*ry fit 32 bits and each second-order moment is one 32x32->64 product.
ATQ_DEV atq_line_fit_point linefit_from_terms(uint32_t W, uint32_t rx, uint32_t ry) {
    const uint32_t wx = W * rx, wy = W * ry;
    return atq_line_fit_point{(int64_t)wx,
                              (int64_t)wy,
                              (int64_t)((uint64_t)wx * rx),
                              (int64_t)((uint64_t)wy * ry),
                              (int64_t)((uint64_t)wx * ry),
                              (int64_t)W,
                              0,
                              0};
}

ATQ_DEV atq_line_fit_point initial_linefit(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                                           const uint8_t *__restrict__ grey) {
    uint32_t W, rx, ry;
    linefit_terms(px, py, w, h, grey, &W, &rx, &ry);
    return linefit_from_terms(W, rx, ry);
}

}  // namespace atq
