// diff.hip - differentiable line-fit / corner stage (forward + backward).
//
// A quad is four lines fitted to the cyclic point ranges [i0,i1], [i1,i2],
// [i2,i3], [i3,i0] of one angle-sorted cluster (the ranges do_indexing forms,
// src/fit_quad.cpp:445-493); corner k is the intersection of lines k and
// k+1 mod 4 (process_fitted_quads, src/main.cpp:1059-1096).
//
// Unlike the reference's integer weights (which read the row from rx,
// include/line_fit_point.h:57), point i has the float weight
//     w_i = s_i * sqrt(1 + gx^2 + gy^2)
// with (gx, gy) the central differences of the image at the point's pixel
// (x+1, y) and s_i a per-point boundary weight, so corners are smooth in the
// pixels and in s. Point i sits at (X, Y) = (x + 1, y + 0.5) (the
// reference's rx/2, ry/2).
//
// Line: weighted centroid m and major-axis angle phi = atan2(2 Cxy, Cxx - Cyy)/2
// of the weighted covariance, direction d = (cos phi, sin phi). Moments are
// accumulated in float32 with Kahan compensation relative to the line's first
// point (keeps the covariance free of cancellation).
// Corner: c = m_a + t d_a, t = cross(m_b - m_a, d_b) / cross(d_a, d_b).
//
// Backward (vector-Jacobian product) for upstream g = dL/dc:
//   dL/dm_a = g + (g.d_a / D) (-d_by, d_bx),  dL/dm_b = (g.d_a / D) (d_by, -d_bx)
//   dL/dphi_a = t (g.n_a) + (g.d_a) t (d_a.d_b) / D,   n = (-d_y, d_x)
//   dL/dphi_b = (g.d_a) ((m_b - m_a).d_b - t (d_a.d_b)) / D
// and per point, with u = X - mx, v = Y - my, A = Cxx - Cyy, B = 2 Cxy:
//   dm/dw_i = (u, v) / W
//   dphi/dw_i = (A (2uv - B) - B (u^2 - v^2 - A)) / (2 W (A^2 + B^2))
// then dL/ds_i = dL/dw_i * w_i / s_i and dL/dI through the four pixels of
// (gx, gy) (atomically accumulated).
#include "kernels.hpp"

namespace atq {

namespace {

struct LineParams {
    float W, mx, my, A, B, phi, dx, dy;
};

ATQ_DEV void point_xy(const atq_cluster_point &p, float ox, float oy, float *X, float *Y) {
    *X = (float)((p.packed_x & COORD_VALUE_MASK) + 1) - ox;
    *Y = ((float)(p.packed_y & COORD_VALUE_MASK) + 0.5f) - oy;
}

// weight and its pixel gradient factors
ATQ_DEV float point_weight(const atq_cluster_point &p, const float *__restrict__ img, uint32_t w,
                           uint32_t h, float s, float *gx, float *gy, float *r, int *px, int *py) {
    const int x = (int)(p.packed_x & COORD_VALUE_MASK) + 1;
    const int y = (int)(p.packed_y & COORD_VALUE_MASK);
    *px = x;
    *py = y;
    float a = 0.0f, b = 0.0f;
    if (x > 0 && x + 1 < (int)w && y > 0 && y + 1 < (int)h) {
        a = img[(size_t)y * w + x + 1] - img[(size_t)y * w + x - 1];
        b = img[(size_t)(y + 1) * w + x] - img[(size_t)(y - 1) * w + x];
    }
    *gx = a;
    *gy = b;
    *r = sqrtf(1.0f + a * a + b * b);
    return s * *r;
}

struct Kahan {
    float s = 0.0f, c = 0.0f;
    ATQ_DEV void add(float v) {
        float y = v - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

ATQ_DEV float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Walks the cyclic range [i0, i1] (inclusive; i1 < i0 wraps) of the cluster.
ATQ_DEV uint32_t range_len(uint32_t i0, uint32_t i1, uint32_t n) {
    return i0 <= i1 ? i1 - i0 + 1 : n - i0 + i1 + 1;
}

// Moments are taken relative to the line's first point (no cancellation in
// the covariance), the centroid is reported in the quad's frame.
ATQ_DEV LineParams fit_line_f(const atq_cluster_point *__restrict__ pts, uint32_t n, uint32_t i0,
                              uint32_t i1, const float *__restrict__ img, uint32_t w, uint32_t h,
                              const float *__restrict__ sw, uint32_t base, float ox, float oy) {
    const uint32_t len = range_len(i0, i1, n);
    float lx, ly;
    point_xy(pts[i0], ox, oy, &lx, &ly);
    Kahan kW, kx, ky, kxx, kyy, kxy;
    for (uint32_t j = __lane_id(); j < len; j += 64) {
        const uint32_t idx = (i0 + j) % n;
        const atq_cluster_point p = pts[idx];
        float gx, gy, r;
        int px, py;
        const float s = sw ? sw[base + idx] : 1.0f;
        const float wt = point_weight(p, img, w, h, s, &gx, &gy, &r, &px, &py);
        float X, Y;
        point_xy(p, ox + lx, oy + ly, &X, &Y);
        kW.add(wt);
        kx.add(wt * X);
        ky.add(wt * Y);
        kxx.add(wt * X * X);
        kyy.add(wt * Y * Y);
        kxy.add(wt * X * Y);
    }
    LineParams L;
    const float W = wave_sum(kW.s - kW.c), Sx = wave_sum(kx.s - kx.c), Sy = wave_sum(ky.s - ky.c);
    const float Sxx = wave_sum(kxx.s - kxx.c), Syy = wave_sum(kyy.s - kyy.c), Sxy = wave_sum(kxy.s - kxy.c);
    const float mx = Sx / W, my = Sy / W;
    L.W = W;
    L.mx = lx + mx;
    L.my = ly + my;
    const float cxx = Sxx / W - mx * mx;
    const float cyy = Syy / W - my * my;
    const float cxy = Sxy / W - mx * my;
    L.A = cxx - cyy;
    L.B = 2.0f * cxy;
    L.phi = 0.5f * atan2f(L.B, L.A);
    L.dx = cosf(L.phi);
    L.dy = sinf(L.phi);
    return L;
}

ATQ_DEV void corner_of(const LineParams &a, const LineParams &b, float *cx, float *cy, float *t,
                       float *D) {
    const float Dx = b.mx - a.mx, Dy = b.my - a.my;
    *D = a.dx * b.dy - a.dy * b.dx;
    *t = (Dx * b.dy - Dy * b.dx) / *D;
    *cx = a.mx + *t * a.dx;
    *cy = a.my + *t * a.dy;
}

}  // namespace

// One wave per quad.
__global__ void __launch_bounds__(256)
k_diff_forward(const float *__restrict__ img, uint32_t w, uint32_t h,
               const atq_cluster_point *__restrict__ pts, const float *__restrict__ sw,
               const atq_diff_quad *__restrict__ quads, uint32_t nq, float *__restrict__ corners,
               float *__restrict__ w_out) {
    const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (q >= nq) return;
    const atq_diff_quad Q = quads[q];
    const atq_cluster_point *cp = pts + Q.cluster_start;
    float ox, oy;
    point_xy(cp[Q.idx[0]], 0.0f, 0.0f, &ox, &oy);
    LineParams L[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        L[k] = fit_line_f(cp, Q.cluster_count, Q.idx[k], Q.idx[(k + 1) & 3], img, w, h, sw,
                          Q.cluster_start, ox, oy);
    const uint32_t lane = __lane_id();
    if (lane < 4) {
        float cx, cy, t, D;
        corner_of(L[lane], L[(lane + 1) & 3], &cx, &cy, &t, &D);
        corners[(size_t)q * 8 + 2 * lane] = cx + ox;
        corners[(size_t)q * 8 + 2 * lane + 1] = cy + oy;
    }
    if (w_out) {
        for (uint32_t j = lane; j < Q.cluster_count; j += 64) {
            float gx, gy, r;
            int px, py;
            const float s = sw ? sw[Q.cluster_start + j] : 1.0f;
            w_out[Q.cluster_start + j] = point_weight(cp[j], img, w, h, s, &gx, &gy, &r, &px, &py);
        }
    }
}

__global__ void __launch_bounds__(256)
k_diff_backward(const float *__restrict__ img, uint32_t w, uint32_t h,
                const atq_cluster_point *__restrict__ pts, const float *__restrict__ sw,
                const atq_diff_quad *__restrict__ quads, uint32_t nq,
                const float *__restrict__ gcorners, float *__restrict__ gw,
                float *__restrict__ gimg) {
    const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (q >= nq) return;
    const atq_diff_quad Q = quads[q];
    const atq_cluster_point *cp = pts + Q.cluster_start;
    const uint32_t n = Q.cluster_count;
    float ox, oy;
    point_xy(cp[Q.idx[0]], 0.0f, 0.0f, &ox, &oy);
    LineParams L[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        L[k] = fit_line_f(cp, n, Q.idx[k], Q.idx[(k + 1) & 3], img, w, h, sw, Q.cluster_start, ox, oy);
    // upstream per line: (d mx, d my, d phi); every lane computes all four
    float gmx[4] = {0, 0, 0, 0}, gmy[4] = {0, 0, 0, 0}, gph[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int a = k, b = (k + 1) & 3;
        const float gx = gcorners[(size_t)q * 8 + 2 * k], gy = gcorners[(size_t)q * 8 + 2 * k + 1];
        float cx, cy, t, D;
        corner_of(L[a], L[b], &cx, &cy, &t, &D);
        const float G1 = gx * L[a].dx + gy * L[a].dy;
        const float dadb = L[a].dx * L[b].dx + L[a].dy * L[b].dy;
        const float Dx = L[b].mx - L[a].mx, Dy = L[b].my - L[a].my;
        gmx[a] += gx + G1 * (-L[b].dy) / D;
        gmy[a] += gy + G1 * (L[b].dx) / D;
        gmx[b] += G1 * L[b].dy / D;
        gmy[b] += G1 * (-L[b].dx) / D;
        const float gna = gx * (-L[a].dy) + gy * L[a].dx;
        gph[a] += t * gna + G1 * t * dadb / D;
        gph[b] += G1 * ((Dx * L[b].dx + Dy * L[b].dy) - t * dadb) / D;
    }
    // per point of each line
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const LineParams &Lk = L[k];
        const uint32_t i0 = Q.idx[k], i1 = Q.idx[(k + 1) & 3];
        const uint32_t len = range_len(i0, i1, n);
        const float AB = Lk.A * Lk.A + Lk.B * Lk.B;
        for (uint32_t j = __lane_id(); j < len; j += 64) {
            const uint32_t idx = (i0 + j) % n;
            const atq_cluster_point p = cp[idx];
            float X, Y;
            point_xy(p, ox, oy, &X, &Y);
            const float u = X - Lk.mx, v = Y - Lk.my;
            const float dA = (u * u - v * v - Lk.A) / Lk.W;
            const float dB = (2.0f * u * v - Lk.B) / Lk.W;
            const float dphi = 0.5f * (Lk.A * dB - Lk.B * dA) / AB;
            const float dLdw = gmx[k] * u / Lk.W + gmy[k] * v / Lk.W + gph[k] * dphi;
            float gx, gy, r;
            int px, py;
            const float s = sw ? sw[Q.cluster_start + idx] : 1.0f;
            point_weight(p, img, w, h, s, &gx, &gy, &r, &px, &py);
            if (gw) atomicAdd(&gw[Q.cluster_start + idx], dLdw * r);
            if (gimg && px > 0 && px + 1 < (int)w && py > 0 && py + 1 < (int)h) {
                const float gdx = dLdw * s * gx / r, gdy = dLdw * s * gy / r;
                atomicAdd(&gimg[(size_t)py * w + px + 1], gdx);
                atomicAdd(&gimg[(size_t)py * w + px - 1], -gdx);
                atomicAdd(&gimg[(size_t)(py + 1) * w + px], gdy);
                atomicAdd(&gimg[(size_t)(py - 1) * w + px], -gdy);
            }
        }
    }
}

int launch_diff_forward(hipStream_t s, const float *img, size_t w, size_t h,
                        const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                        size_t nq, float *corners, float *w_out) {
    if (nq == 0) return ATQ_OK;
    k_diff_forward<<<atq_div_up(nq * 64, 256), 256, 0, s>>>(img, (uint32_t)w, (uint32_t)h, pts, sw,
                                                             quads, (uint32_t)nq, corners, w_out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_diff_backward(hipStream_t s, const float *img, size_t w, size_t h,
                         const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                         size_t nq, const float *gcorners, float *gw, float *gimg) {
    if (nq == 0) return ATQ_OK;
    k_diff_backward<<<atq_div_up(nq * 64, 256), 256, 0, s>>>(img, (uint32_t)w, (uint32_t)h, pts,
                                                              sw, quads, (uint32_t)nq, gcorners, gw,
                                                              gimg);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
