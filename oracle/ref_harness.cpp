// ref_harness.cpp - exposes the reference's SYCL-free header functions
// (include/boundary_point.h, include/cluster_bounds.h,
// include/line_fit_point.h, include/combinadics.h of
// carbotaniuman/sycl-apriltag-test) through a C ABI so that tests can check
// the oracle against the reference's own code.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/ with
// the reference's include directory on the include path; nothing of the
// reference is copied here. Only tests/ load the result.
#include <cstdint>
#include <cstring>

#include "boundary_point.h"
#include "cluster_bounds.h"
#include "combinadics.h"
#include "line_fit_point.h"

using RefCombinadics = Combinadics<uint8_t, 15, 4, 16>;  // src/fit_quad.cpp:391

extern "C" {

float ref_calc_theta(uint16_t px, uint16_t py, float cx, float cy) {
    BoundaryPoint p{px, py};
    return p.calc_theta(cx, cy);
}

void ref_point_fields(uint16_t px, uint16_t py, int32_t *out6) {
    BoundaryPoint p{px, py};
    out6[0] = p.x_with_dx();
    out6[1] = p.y_with_dy();
    out6[2] = p.dx();
    out6[3] = p.dy();
    out6[4] = p.gx();
    out6[5] = p.gy();
}

void ref_bounds_from_point(uint16_t px, uint16_t py, uint32_t start, void *out) {
    ClusterBounds b = ClusterBounds::inital_from_point(BoundaryPoint{px, py}, start);
    std::memcpy(out, &b, sizeof(b));
}

void ref_reduce_bounds(const void *a, const void *b, void *out) {
    ClusterBounds x, y;
    std::memcpy(&x, a, sizeof(x));
    std::memcpy(&y, b, sizeof(y));
    ClusterBounds r = reduce_bounds(x, y);
    std::memcpy(out, &r, sizeof(r));
}

void ref_bounds_cx_cy_dot(const void *b, float *out3) {
    ClusterBounds c;
    std::memcpy(&c, b, sizeof(c));
    out3[0] = c.cx();
    out3[1] = c.cy();
    out3[2] = c.dot();
}

int ref_valid_blob(const void *b, int64_t w, int64_t h) {
    ClusterBounds c;
    std::memcpy(&c, b, sizeof(c));
    return ValidBlobFilter((size_t)w, (size_t)h)(c) ? 1 : 0;
}

void ref_initial_linefit(uint16_t px, uint16_t py, uint32_t w, uint32_t h,
                         uint8_t *grey, void *out) {
    ClusterPoint p{px, py, 0.0f};
    LineFitPoint r = compute_initial_linefit(p, w, h, grey);
    std::memcpy(out, &r, sizeof(r));
}

int ref_n_choose_k(int n, int k) { return RefCombinadics::n_choose_k(n, k); }

void ref_combinadics_decode(int num, int64_t *out4) {
    auto r = RefCombinadics::decode((uint8_t)num);
    for (int i = 0; i < 4; i++) out4[i] = (int64_t)r[i];
}

int ref_sizeof(int which) {
    switch (which) {
    case 0: return sizeof(BoundaryPoint);
    case 1: return sizeof(ClusterPoint);
    case 2: return sizeof(ClusterBounds);
    case 3: return sizeof(LineFitPoint);
    case 4: return sizeof(ClusterExtents);
    default: return -1;
    }
}
}
