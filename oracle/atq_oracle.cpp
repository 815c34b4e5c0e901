// atq_oracle.cpp - CPU restatement of the reference AprilTag quad front-end.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle and the CPU
// baseline: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it, and only as the checker. The product path never calls it.
//
// Each function restates one reference stage and cites the reference file
// and lines it follows (paths relative to carbotaniuman/sycl-apriltag-test).
// Kernels are restated literally - the same index spaces, the same
// work-group emulation where the reference's result depends on it - so that
// the GPU implementation, which is organised differently, is checked against
// an independent statement of the reference.
//
// Canonical choices where the reference is not deterministic (each is
// exercised by tests/test_oracle_*.py):
//   * every oneDPL sort is stable (src/main.cpp:515, 737, 857); compaction
//     keeps input order;
//   * floating point is IEEE binary32/64 with no FMA contraction (built with
//     -ffp-contract=off); hypotf(a,b) is (float)sqrt((double)a*a+(double)b*b),
//     which is what glibc's hypotf computes;
//   * moment prefix sums are exact int64 (the reference adds int64 bit
//     patterns as doubles, exact while sums stay below 2^53);
//   * fit_line's int64 products wrap modulo 2^64;
//   * cluster indices are u32 (the reference's u16 caps at 65535);
//   * do_indexing reads of the moment array past its end (possible only
//     because do_indexing pairs peak segment e with cluster extents[e]) read
//     zeros and are counted.
//
// Parity status per function is recorded in the header table of
// oracle/README.md.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <parallel/algorithm>

namespace {

// ---------------------------------------------------------------- records
struct Extrema { uint8_t mn, mx; };

// include/boundary_point.h:8-36
enum HalfPixel : uint16_t { RIGHT = 0, BOTTOM_LEFT = 1, BOTTOM = 2, BOTTOM_RIGHT = 3 };
// (1<<15) & (1<<14) == 0: the half-pixel is never extracted.
constexpr uint16_t COORDINATE_HALF_PIXEL_MASK = (1u << 15) & (1u << 14);
constexpr uint16_t COORDINATE_COLOR_DIRECTION_MASK = 1u << 15;
constexpr uint16_t COORDINATE_VALUE_MASK = (1u << 14) - 1;
constexpr uint32_t LABEL_PIXEL_MASK = 0x80000000u;  // include/segmentation.h:7
constexpr uint32_t LABEL_VALUE_MASK = 0x7FFFFFFFu;  // include/segmentation.h:9

// Fixed mode (`fx`) reads both bits, as pack_half_pixel (include/boundary_point.h:26-28)
// stores them: the plane the record came from.
constexpr uint16_t FIXED_HALF_PIXEL_MASK = (1u << 15) | (1u << 14);

inline uint16_t extract_half_pixel(uint16_t c, bool fx = false) {
    return (uint16_t)((c & (fx ? FIXED_HALF_PIXEL_MASK : COORDINATE_HALF_PIXEL_MASK)) >> 14);
}
// include/boundary_point.h:67-95
inline int dx_of(uint16_t px, bool fx = false) {
    switch (extract_half_pixel(px, fx)) {
    case RIGHT: return 1; case BOTTOM_LEFT: return -1;
    case BOTTOM: return 0; case BOTTOM_RIGHT: return 1; default: return 0;
    }
}
inline int dy_of(uint16_t px, bool fx = false) {
    switch (extract_half_pixel(px, fx)) {
    case RIGHT: return 0; case BOTTOM_LEFT: return 1;
    case BOTTOM: return 1; case BOTTOM_RIGHT: return 1; default: return 0;
    }
}

struct BoundaryPoint {
    uint16_t packed_x, packed_y;
    bool b2w() const { return packed_y & COORDINATE_COLOR_DIRECTION_MASK; }
    uint16_t xv() const { return packed_x & COORDINATE_VALUE_MASK; }
    uint16_t yv() const { return packed_y & COORDINATE_VALUE_MASK; }
    uint16_t x_with_dx(bool fx = false) const { return (uint16_t)(2 * xv() + dx_of(packed_x, fx)); }
    uint16_t y_with_dy(bool fx = false) const { return (uint16_t)(2 * yv() + dy_of(packed_x, fx)); }
    int gx(bool fx = false) const { return b2w() ? dx_of(packed_x, fx) : -dx_of(packed_x, fx); }
    int gy(bool fx = false) const { return b2w() ? dy_of(packed_x, fx) : -dy_of(packed_x, fx); }
    // include/boundary_point.h:48-61
    float calc_theta(float cx, float cy, bool fx = false) const {
        const float quadrants[2][2] = {{-1.0f * (1 << 16), (float)(1 << 17)},
                                       {0.0f, (float)(1 << 16)}};
        float dx = (float)x_with_dx(fx) - cx;
        float dy = (float)y_with_dy(fx) - cy;
        float q = quadrants[dx > 0][dy > 0];
        if ((dx > 0) == (dy > 0)) return q + dy / dx;
        return q - dx / dy;
    }
};
static_assert(sizeof(BoundaryPoint) == 4, "");

struct ClusterPoint { uint16_t packed_x, packed_y; float slope; };
static_assert(sizeof(ClusterPoint) == 8, "");

struct ClusterExtents { uint32_t start, count; };

// include/cluster_bounds.h:14-83
struct ClusterBounds {
    uint16_t x_min, x_max, y_min, y_max;
    int32_t gx_sum, gy_sum;
    int64_t pxgx_pygy_sum;
    uint32_t start, count;

    static ClusterBounds from_point(BoundaryPoint p, uint32_t start, bool fx = false) {
        ClusterBounds b;
        b.x_min = b.x_max = p.x_with_dx(fx);
        b.y_min = b.y_max = p.y_with_dy(fx);
        b.gx_sum = p.gx(fx);
        b.gy_sum = p.gy(fx);
        b.pxgx_pygy_sum =
            (int64_t)p.x_with_dx(fx) * p.gx(fx) + (int64_t)p.y_with_dy(fx) * p.gy(fx);
        b.start = start;
        b.count = 1;
        return b;
    }
    // include/cluster_bounds.h:47-49: float*0.5f, then a double add, then
    // rounded back to float.
    float cx() const { return (float)((double)((float)(x_min + x_max) * 0.5f) + 0.05118); }
    float cy() const { return (float)((double)((float)(y_min + y_max) * 0.5f) + -0.028581); }
    // include/cluster_bounds.h:51-66
    float dot() const {
        float cx_cy = (float)((double)((x_min + x_max) * gx_sum + (y_min + y_max) * gy_sum) * 0.5);
        float gxc = (float)(0.05118 * (double)(float)gx_sum);
        float gyc = (float)(-0.028581 * (double)(float)gy_sum);
        float r = 2.0f * (float)pxgx_pygy_sum;
        r = r - cx_cy;
        r = r - gxc;
        r = r - gyc;
        return r;
    }
};
static_assert(sizeof(ClusterBounds) == 32, "");

inline ClusterBounds reduce_bounds(const ClusterBounds &a, const ClusterBounds &b) {
    ClusterBounds o;
    o.x_min = std::min(a.x_min, b.x_min);
    o.x_max = std::max(a.x_max, b.x_max);
    o.y_min = std::min(a.y_min, b.y_min);
    o.y_max = std::max(a.y_max, b.y_max);
    o.start = std::min(a.start, b.start);
    o.count = a.count + b.count;
    o.pxgx_pygy_sum = a.pxgx_pygy_sum + b.pxgx_pygy_sum;
    o.gx_sum = a.gx_sum + b.gx_sum;
    o.gy_sum = a.gy_sum + b.gy_sum;
    return o;
}

// include/cluster_bounds.h:85-120
struct ValidBlobFilter {
    size_t tag_width = 4;
    bool normal_border = true;
    bool reversed_border = false;
    size_t min_cluster_pixels = 5;
    size_t max_cluster_pixels;
    ValidBlobFilter(size_t w, size_t h) : max_cluster_pixels(2 * (2 * w + 2 * h)) {}
    bool operator()(const ClusterBounds &b) const {
        if (b.count < min_cluster_pixels) return false;
        if (b.count < 24) return false;
        if (b.count > max_cluster_pixels) return false;
        // int product compared with a size_t: the product is never negative.
        if ((size_t)((b.x_max - b.x_min) * (b.y_max - b.y_min)) < tag_width) return false;
        bool quad_reversed = b.dot() < 0.0;
        if (!reversed_border && quad_reversed) return false;
        if (!normal_border && !quad_reversed) return false;
        return true;
    }
};

// include/line_fit_point.h:11-48
struct LineFitPoint {
    int64_t Mx, My, Mxx, Myy, Mxy, W, a, b;
};
static_assert(sizeof(LineFitPoint) == 64, "");
inline LineFitPoint lf_add(LineFitPoint l, const LineFitPoint &r) {
    l.Mx += r.Mx; l.My += r.My; l.Mxx += r.Mxx; l.Myy += r.Myy; l.Mxy += r.Mxy; l.W += r.W;
    return l;
}
inline LineFitPoint lf_sub(LineFitPoint l, const LineFitPoint &r) {
    l.Mx -= r.Mx; l.My -= r.My; l.Mxx -= r.Mxx; l.Myy -= r.Myy; l.Mxy -= r.Mxy; l.W -= r.W;
    return l;
}

// glibc's hypotf: the double-precision root of the exact sum, rounded once.
inline float hypot_f(float a, float b) {
    double da = a, db = b;
    return (float)std::sqrt(da * da + db * db);
}

// include/line_fit_point.h:50-86 (with its `y = rx / 2`; fixed mode: `ry / 2`)
inline LineFitPoint compute_initial_linefit(ClusterPoint p, uint32_t width, uint32_t height,
                                            const uint8_t *grey, bool fx = false) {
    BoundaryPoint bp{p.packed_x, p.packed_y};
    int64_t rx = (int64_t)bp.x_with_dx(fx) + 1;
    int64_t ry = (int64_t)bp.y_with_dy(fx) + 1;
    uint32_t x = (uint32_t)(rx / 2);
    uint32_t y = (uint32_t)((fx ? ry : rx) / 2);
    int64_t W = 1;
    if (x > 0 && x + 1 < width && y > 0 && y + 1 < height) {
        int gx = (int)grey[(size_t)y * width + x + 1] - (int)grey[(size_t)y * width + x - 1];
        int gy = (int)grey[(size_t)(y + 1) * width + x] - (int)grey[(size_t)(y - 1) * width + x];
        W += (int64_t)hypot_f((float)gx, (float)gy);
    }
    LineFitPoint r{};
    r.Mx = W * rx; r.My = W * ry;
    r.Mxx = W * rx * rx; r.Myy = W * ry * ry; r.Mxy = W * rx * ry;
    r.W = W;
    return r;
}

struct PeakExtents { uint32_t start, count; };
struct Corner { uint32_t line_fit_point_index; uint32_t cluster_index; float error; };
static_assert(sizeof(Corner) == 12, "");
struct FittedQuad {
    LineFitPoint moments[4];
    uint16_t num_in_moments[4];
    uint16_t indices[4];
};
static_assert(sizeof(FittedQuad) == 272, "");
struct QuadCorners {
    float corners[4][2];
    uint8_t reversed_border;
    uint8_t pad_[3];
    uint32_t blob_index;
};
static_assert(sizeof(QuadCorners) == 40, "");

// --------------------------------------------------------- fit_quad core
// src/fit_quad.cpp:15-35. `bound` is the number of valid entries readable
// at `pts` (reads at or past it return zeros and bump *oob).
inline LineFitPoint pt_at(const LineFitPoint *pts, int64_t i, int64_t bound, uint64_t *oob) {
    if (i < 0 || i >= bound) {
        if (oob) ++*oob;
        return LineFitPoint{};
    }
    return pts[i];
}
inline LineFitPoint get_moment(const LineFitPoint *pts, size_t n, size_t i0, size_t i1,
                               size_t *total, int64_t bound, uint64_t *oob) {
    LineFitPoint r{};
    *total = 0;
    if (i0 < i1) {
        *total = i1 - i0 + 1;
        r = pt_at(pts, (int64_t)i1, bound, oob);
        if (i0 > 0) r = lf_sub(r, pt_at(pts, (int64_t)i0 - 1, bound, oob));
    } else if (i1 < i0) {
        *total = n - i0 + i1 + 1;
        r = lf_add(lf_sub(pt_at(pts, (int64_t)n - 1, bound, oob),
                          pt_at(pts, (int64_t)i0 - 1, bound, oob)),
                   pt_at(pts, (int64_t)i1, bound, oob));
    }
    return r;
}

inline int64_t wmul(int64_t a, int64_t b) { return (int64_t)((uint64_t)a * (uint64_t)b); }
inline int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }

// src/fit_quad.cpp:37-105
inline void fit_line(const LineFitPoint &m, size_t num, float *line, float *err, float *mse) {
    int64_t Csxx = wsub(wmul(m.Mxx, m.W), wmul(m.Mx, m.Mx));
    int64_t Csxy = wsub(wmul(m.Mxy, m.W), wmul(m.Mx, m.My));
    int64_t Csyy = wsub(wmul(m.Myy, m.W), wmul(m.My, m.My));
    int64_t corr = wmul(wmul(m.W, m.W), 4);
    float sdist = hypot_f((float)wsub(Csxx, Csyy), (float)wmul(2, Csxy));
    float num_f = (float)wadd(Csxx, Csyy) - sdist;
    float eig_err = num_f / (float)wmul(corr, 2);
    if (line) {
        float nx1 = (float)wsub(Csxx, Csyy) - sdist;
        float ny1 = (float)wmul(2, Csxy);
        float nx2 = (float)wmul(2, Csxy);
        float ny2 = (float)wsub(Csyy, Csxx) - sdist;
        float a1 = nx1 * nx1, b1 = ny1 * ny1;
        float M1 = a1 + b1;
        float a2 = nx2 * nx2, b2 = ny2 * ny2;
        float M2 = a2 + b2;
        float nx, ny;
        if (M1 > M2) { nx = nx1; ny = ny1; } else { nx = nx2; ny = ny2; }
        line[0] = (float)m.Mx / (float)wmul(m.W, 2);
        line[1] = (float)m.My / (float)wmul(m.W, 2);
        float length = hypot_f(nx, ny);
        if ((double)std::fabs(length) < 1e-12) {
            line[2] = 0; line[3] = 0;
        } else {
            line[2] = nx / length;
            line[3] = ny / length;
        }
    }
    if (err) *err = (float)num * eig_err;
    if (mse) *mse = eig_err;
}

// src/fit_quad.cpp:10-13: (float)exp(-(j-3)^2/2)
const float FILTER_DATA[7] = {0x1.6c0504p-7f, 0x1.152aaap-3f, 0x1.368b3p-1f, 0x1p+0f,
                              0x1.368b3p-1f, 0x1.152aaap-3f, 0x1.6c0504p-7f};

// include/combinadics.h:14-60 instantiated as Combinadics<uint8_t, 15, 4, 16>
// (src/fit_quad.cpp:391): binomials stored in uint8_t (they wrap for n > 10)
// and decode by std::upper_bound over each row.
struct Combinadics {
    uint8_t tab[5][16];
    Combinadics() {
        std::memset(tab, 0, sizeof(tab));
        for (int n = 0; n <= 15; n++) tab[0][n] = 1;
        for (int k = 1; k <= 4; k++)
            for (int n = 1; n <= 15; n++) tab[k][n] = (uint8_t)(tab[k - 1][n - 1] + tab[k][n - 1]);
    }
    uint8_t n_choose_k(size_t n, size_t k) const { return tab[k][n]; }
    void decode(uint8_t num, size_t out[4]) const {
        for (int i = 0; i < 4; i++) out[i] = 0;
        for (int i = 0; i < 4; i++) {
            const uint8_t *cur = tab[4 - i];
            const uint8_t *found = std::upper_bound(cur, cur + 16, num);
            if (found == cur) return;
            --found;
            num = (uint8_t)(num - *found);
            out[i] = (size_t)(found - cur);
        }
    }
};
const Combinadics &combinadics() {
    static Combinadics c;
    return c;
}

int nthreads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // namespace

// =======================================================================
extern "C" {

struct orc_stats {
    int64_t width, height;
    int64_t n_sparse_points;   // compacted boundary records
    int64_t n_pairs;           // cluster-pair segments
    int64_t n_clusters;        // clusters passing ValidBlobFilter
    int64_t n_cluster_points;  // points of those clusters
    int64_t n_corners;         // corners with error != 0
    int64_t n_segments;        // peak segments
    int64_t n_quads;           // process_fitted_quads survivors
    int64_t oob_reads;         // do_indexing reads past the moment array
    int64_t n_fitted;          // do_indexing outputs with a chosen combo
};

// Decimation: upstream apriltag image_u8_decimate for integer factors: point
// subsample, output 1+(w-1)/f by 1+(h-1)/f (not in the reference, notes.txt:18-21).
void orc_decimate(const uint8_t *in, int64_t w_in, int64_t h_in, int f, uint8_t *out) {
    int64_t w = 1 + (w_in - 1) / f, h = 1 + (h_in - 1) / f;
#pragma omp parallel for
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) out[y * w + x] = in[(y * f) * w_in + x * f];
}

// src/threshold.cpp:8-95 (three kernels)
void orc_threshold(const uint8_t *g, int64_t w, int64_t h, uint8_t *extrema, uint8_t *filtered,
                   uint8_t *thr) {
    const int64_t tw = w / 4, th = h / 4;
    Extrema *ext = reinterpret_cast<Extrema *>(extrema);
    Extrema *fil = reinterpret_cast<Extrema *>(filtered);
    // K1: src/threshold.cpp:8-33
#pragma omp parallel for
    for (int64_t ty = 0; ty < th; ty++)
        for (int64_t tx = 0; tx < tw; tx++) {
            uint8_t mn = 255, mx = 0;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    uint8_t v = g[(ty * 4 + i) * (tw * 4) + (tx * 4 + j)];
                    mn = std::min(mn, v);
                    mx = std::max(mx, v);
                }
            ext[ty * tw + tx] = {mn, mx};
        }
    // K2: src/threshold.cpp:35-66 (3x3 tile neighbourhood, clipped at edges)
#pragma omp parallel for
    for (int64_t ty = 0; ty < th; ty++)
        for (int64_t tx = 0; tx < tw; tx++) {
            uint8_t mn = 255, mx = 0;
            for (int i = -1; i <= 1; i++)
                for (int j = -1; j <= 1; j++) {
                    int64_t ey = ty + i, ex = tx + j;
                    if (ex < 0 || ex >= tw || ey < 0 || ey >= th) continue;
                    Extrema v = ext[ey * tw + ex];
                    mn = std::min(mn, v.mn);
                    mx = std::max(mx, v.mx);
                }
            fil[ty * tw + tx] = {mn, mx};
        }
    // K3: src/threshold.cpp:68-95
#pragma omp parallel for
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) {
            Extrema f = fil[(y / 4) * (w / 4) + x / 4];
            uint8_t mn = f.mn, mx = f.mx;
            uint8_t comp = (uint8_t)(mn + (mx - mn) / 2);
            uint8_t v = g[y * w + x];
            uint8_t o;
            if (mx - mn < 5) o = 127;
            else o = v > comp ? 255 : 0;
            thr[y * w + x] = o;
        }
}

// BkeBitmap, src/segmentation.cpp:47-71
enum : uint16_t {
    TL255 = 1 << 0, TR255 = 1 << 1, BL255 = 1 << 2, BR255 = 1 << 3,
    MU_P255 = 1 << 4, MU_Q255 = 1 << 5, MU_R255 = 1 << 6, MU_S255 = 1 << 7,
    TL0 = 1 << 8, TR0 = 1 << 9, BL0 = 1 << 10, BR0 = 1 << 11,
    MU_Q0 = 1 << 13, MU_R0 = 1 << 14, MU_S0 = 1 << 15,
};

// include/union_find.h:19-77, run sequentially.
static uint32_t uf_find(const uint32_t *buf, uint32_t n) {
    while (buf[n] != n) n = buf[n];
    return n;
}
static void uf_find_compress(uint32_t *buf, uint32_t n) {
    uint32_t orig = n;
    while (true) {
        uint32_t next = buf[n];
        if (next == n) break;
        n = next;
        if (n < buf[orig]) buf[orig] = n;
    }
}
static void uf_merge(uint32_t *buf, uint32_t a, uint32_t b) {
    while (true) {
        a = uf_find(buf, a);
        b = uf_find(buf, b);
        if (a < b) {
            uint32_t old = buf[b];
            if (a < old) buf[b] = a;
            if (old == b) break;
            b = old;
        } else if (b < a) {
            uint32_t old = buf[a];
            if (b < old) buf[a] = b;
            if (old == a) break;
            a = old;
        } else {
            break;
        }
    }
}

// src/segmentation.cpp:95-392 (BKE: INIT, COMPRESS, MERGE, COMPRESS,
// FINAL_LABELLING). `sizes` (w*h) is zeroed here: the reference memsets only
// 2^16 entries (src/main.cpp:310-311) and parity assumes the whole array.
void orc_segmentation(const uint8_t *thr, int64_t w, int64_t h, uint32_t *scratch,
                      uint32_t *labels, uint32_t *sizes) {
    const int64_t bw = w / 2, bh = h / 2;
    std::memset(sizes, 0, sizeof(uint32_t) * w * h);
    // INIT: src/segmentation.cpp:101-262
#pragma omp parallel for
    for (int64_t by = 0; by < bh; by++)
        for (int64_t bx = 0; bx < bw; bx++) {
            const int64_t x = bx * 2, y = by * 2, id = y * w + x;
            uint16_t info = 0;
            uint32_t check = 0;
            auto classify = [&](int64_t idx, uint32_t mask, uint16_t f255, uint16_t f0) {
                if (thr[idx] == 255) { check |= mask; info |= f255; }
                else if (thr[idx] == 0) { check |= mask; info |= f0; }
            };
            classify(id, 0x777u, TL255, TL0);
            classify(id + 1, 0x777u << 1, TR255, TR0);
            classify(id + w, 0x777u << 4, BL255, BL0);
            if (thr[id + w + 1] == 255) info |= BR255;
            else if (thr[id + w + 1] == 0) info |= BR0;
            if (x == 0) check &= 0xEEEE;
            else if (x + 2 >= w) check &= 0x7777;
            if (y == 0) check &= 0xFFF0;
            else if (y + 2 >= h) check &= 0x0FFF;
            int64_t off255 = 0, off0 = 1;
            auto test = [&](int dy, int dx, int bit, uint16_t extra) {
                if (!(check & (1u << bit))) return;
                uint8_t v = thr[id + dy * w + dx];
                auto offset_for = [&](uint16_t e) -> int64_t {
                    switch (e) {
                    case MU_P255: return -2 * w - 2;
                    case MU_Q255: return -2 * w;
                    case MU_R255: return -2 * w + 2;
                    case MU_S255: return -2;
                    default: return 0;
                    }
                };
                if (v == 255) {
                    if (off255 == 0) off255 = offset_for(extra);
                    else info |= extra;
                } else if (v == 0) {
                    if (off0 == 1) off0 = offset_for(extra) + 1;
                    else info |= (uint16_t)(extra << 8);
                }
            };
            test(-1, -1, 0, MU_P255);
            test(-1, 0, 1, MU_Q255);
            test(-1, 1, 2, MU_Q255);
            test(-1, 2, 3, MU_R255);
            test(0, -1, 4, MU_S255);
            test(1, -1, 8, MU_S255);
            scratch[id] = (uint32_t)(id + off255);
            scratch[id + 1] = (uint32_t)(id + off0);
            scratch[id + w] = info;
        }
    // pre-merge COMPRESS, src/segmentation.cpp:265-266 (sequential here)
    for (int64_t by = 0; by < bh; by++)
        for (int64_t bx = 0; bx < bw; bx++) {
            uint32_t id = (uint32_t)(by * 2 * w + bx * 2);
            uf_find_compress(scratch, id);
            uf_find_compress(scratch, id + 1);
        }
    // MERGE: src/segmentation.cpp:269-306
    for (int64_t by = 0; by < bh; by++)
        for (int64_t bx = 0; bx < bw; bx++) {
            uint32_t id = (uint32_t)(by * 2 * w + bx * 2);
            uint16_t info = (uint16_t)scratch[id + w];
            uint32_t W2 = (uint32_t)(2 * w);
            if (info & MU_Q255) uf_merge(scratch, id, id - W2);
            if (info & MU_R255) uf_merge(scratch, id, id - W2 + 2);
            if (info & MU_S255) uf_merge(scratch, id, id - 2);
            if (info & MU_Q0) uf_merge(scratch, id + 1, id - W2 + 1);
            if (info & MU_R0) uf_merge(scratch, id + 1, id - W2 + 3);
            if (info & MU_S0) uf_merge(scratch, id + 1, id - 1);
        }
    // COMPRESS: src/segmentation.cpp:309-310
    for (int64_t by = 0; by < bh; by++)
        for (int64_t bx = 0; bx < bw; bx++) {
            uint32_t id = (uint32_t)(by * 2 * w + bx * 2);
            uf_find_compress(scratch, id);
            uf_find_compress(scratch, id + 1);
        }
    // FINAL_LABELLING: src/segmentation.cpp:316-389
    for (int64_t by = 0; by < bh; by++)
        for (int64_t bx = 0; bx < bw; bx++) {
            const int64_t id = by * 2 * w + bx * 2;
            uint32_t l255 = scratch[id], l0 = scratch[id + 1];
            uint16_t info = (uint16_t)scratch[id + w];
            uint32_t c255 = (uint32_t)__builtin_popcount(info & 0x000Fu);
            if (c255 > 0) sizes[l255] += c255;
            uint32_t c0 = (uint32_t)__builtin_popcount(info & 0x0F00u);
            if (c0 > 0) sizes[l0] += c0;
            auto lab = [&](uint16_t b255, uint16_t b0) -> uint32_t {
                if (info & b255) return LABEL_PIXEL_MASK | l255;
                if (info & b0) return l0;
                return 0;
            };
            labels[id] = lab(TL255, TL0);
            labels[id + 1] = lab(TR255, TR0);
            labels[id + w] = lab(BL255, BL0);
            labels[id + w + 1] = lab(BR255, BR0);
        }
}

// src/find_boundaries.cpp:24-169, emulating the nd_range of 8x32
// work-groups with the 1 KiB local tile. points/keys hold 4*w*h entries and
// are pre-filled here the way src/main.cpp:312-315 does.
void orc_find_boundaries(const uint32_t *labels, const uint32_t *sizes, int64_t w, int64_t h,
                         uint32_t *points_out, uint64_t *keys) {
    constexpr int64_t BH = 8, BW = 32;
    BoundaryPoint *points = reinterpret_cast<BoundaryPoint *>(points_out);
    const int64_t n = w * h;
    std::memset(points, 0, sizeof(BoundaryPoint) * 4 * n);
    std::memset(keys, 0xFF, sizeof(uint64_t) * 4 * n);
    const int64_t gh = (h + (BH - 1) - 1) / (BH - 1);
    const int64_t gw = (w + (BW - 2) - 1) / (BW - 2);
#pragma omp parallel for
    for (int64_t g0 = 0; g0 < gh; g0++) {
        uint32_t lds[BH][BW];
        for (int64_t g1 = 0; g1 < gw; g1++) {
            for (int64_t ly = 0; ly < BH; ly++)
                for (int64_t lx = 0; lx < BW; lx++) {
                    int64_t x = g1 * (BW - 2) + lx, y = g0 * (BH - 1) + ly;
                    lds[ly][lx] = (x >= w || y >= h) ? 0u : labels[y * w + x];
                }
            for (int64_t ly = 0; ly < BH - 1; ly++)
                for (int64_t lx = 1; lx < BW - 1; lx++) {
                    int64_t x = g1 * (BW - 2) + lx, y = g0 * (BH - 1) + ly;
                    if (x >= w || y >= h) continue;
                    const int64_t lin = y * w + x;
                    uint32_t li = lds[ly][lx];
                    uint32_t local_label = li & LABEL_VALUE_MASK;
                    bool local_small = sizes[local_label] < 25;
                    uint32_t test_local = li & LABEL_PIXEL_MASK;
                    auto handle = [&](int dy, int dx, int64_t plane, uint16_t hp) {
                        uint32_t ti = lds[ly + dy][lx + dx];
                        uint32_t test_other = ti & LABEL_PIXEL_MASK;
                        if (test_other == test_local) return;
                        uint32_t test_label = ti & LABEL_VALUE_MASK;
                        bool test_small = sizes[test_label] < 25;
                        BoundaryPoint p{(uint16_t)((uint16_t)x | (uint16_t)(hp << 14)),
                                        test_other > test_local
                                            ? (uint16_t)(COORDINATE_COLOR_DIRECTION_MASK | y)
                                            : (uint16_t)y};
                        uint64_t key = ((uint64_t)std::min(local_label, test_label) << 32) |
                                       (uint64_t)std::max(local_label, test_label);
                        int64_t oi = n * plane + lin;
                        if (!local_small && !test_small) {
                            points[oi] = p;
                            keys[oi] = key;
                        }
                    };
                    handle(0, 1, 0, RIGHT);
                    {
                        uint32_t left = lds[ly][lx - 1];
                        uint32_t down = lds[ly + 1][lx];
                        bool would_dup = (left & LABEL_PIXEL_MASK) != (down & LABEL_PIXEL_MASK) &&
                                         sizes[left & LABEL_VALUE_MASK] >= 25 &&
                                         sizes[down & LABEL_VALUE_MASK] >= 25;
                        if (!would_dup) handle(1, -1, 1, BOTTOM_LEFT);
                    }
                    handle(1, 0, 2, BOTTOM);
                    handle(1, 1, 3, BOTTOM_RIGHT);
                }
        }
    }
}

// Compaction + key sort: src/main.cpp:467-477 (copy_if on the point's u32 image != 0) and
// src/main.cpp:515-518 (sort_by_key; stable here). Returns N_bp.
int64_t orc_compact_sort(const uint32_t *points_sparse, const uint64_t *keys_sparse,
                         int64_t n_sparse, uint32_t *out_points, uint64_t *out_keys) {
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n_sparse; i++)
        if (points_sparse[i] != 0) idx.push_back(i);
    __gnu_parallel::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        return keys_sparse[a] < keys_sparse[b];
    });
    for (size_t i = 0; i < idx.size(); i++) {
        out_points[i] = points_sparse[idx[i]];
        out_keys[i] = keys_sparse[idx[i]];
    }
    return (int64_t)idx.size();
}

// Cluster bounds: src/main.cpp:571-585 (reduce_by_segment of inital_from_point(p, i)).
static int64_t cluster_bounds_m(const uint32_t *pts_u32, const uint64_t *keys, int64_t n,
                                void *out_v, bool fx) {
    const BoundaryPoint *pts = reinterpret_cast<const BoundaryPoint *>(pts_u32);
    ClusterBounds *out = reinterpret_cast<ClusterBounds *>(out_v);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) {
        ClusterBounds b = ClusterBounds::from_point(pts[i], (uint32_t)i, fx);
        if (i == 0 || keys[i] != keys[i - 1]) out[m++] = b;
        else out[m - 1] = reduce_bounds(out[m - 1], b);
    }
    return m;
}
int64_t orc_cluster_bounds(const uint32_t *pts_u32, const uint64_t *keys, int64_t n,
                           void *out_v) {
    return cluster_bounds_m(pts_u32, keys, n, out_v, false);
}

// Filter: src/main.cpp:604-609 (copy_if with ValidBlobFilter(width, height)).
int64_t orc_filter_bounds(const void *in_v, int64_t n, int64_t w, int64_t h, void *out_v) {
    const ClusterBounds *in = reinterpret_cast<const ClusterBounds *>(in_v);
    ClusterBounds *out = reinterpret_cast<ClusterBounds *>(out_v);
    ValidBlobFilter f((size_t)w, (size_t)h);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (f(in[i])) out[m++] = in[i];
    return m;
}

// Cluster assignment: src/main.cpp:621-675 (lower_bound over the filtered clusters, the
// theta of each point, copy_if on the index sentinel). Returns N_fp.
static int64_t cluster_points_m(const uint32_t *pts_u32, int64_t n, const void *filt_v,
                                int64_t ncl, void *out_pts_v, uint32_t *out_idx, bool fx) {
    const BoundaryPoint *pts = reinterpret_cast<const BoundaryPoint *>(pts_u32);
    const ClusterBounds *filt = reinterpret_cast<const ClusterBounds *>(filt_v);
    ClusterPoint *out = reinterpret_cast<ClusterPoint *>(out_pts_v);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) {
        const ClusterBounds *it = std::lower_bound(
            filt, filt + ncl, (uint32_t)i,
            [](const ClusterBounds &a, uint32_t b) { return (a.start + a.count) <= b; });
        if (it == filt + ncl) continue;
        if ((uint32_t)i < it->start || (uint32_t)i >= it->start + it->count) continue;
        out[m] = ClusterPoint{pts[i].packed_x, pts[i].packed_y,
                              pts[i].calc_theta(it->cx(), it->cy(), fx)};
        out_idx[m] = (uint32_t)(it - filt);
        m++;
    }
    return m;
}
int64_t orc_cluster_points(const uint32_t *pts_u32, int64_t n, const void *filt_v, int64_t ncl,
                           void *out_pts_v, uint32_t *out_idx) {
    return cluster_points_m(pts_u32, n, filt_v, ncl, out_pts_v, out_idx, false);
}

// Extents: src/main.cpp:700-717
void orc_cluster_extents(const void *filt_v, int64_t ncl, void *out_v) {
    const ClusterBounds *filt = reinterpret_cast<const ClusterBounds *>(filt_v);
    ClusterExtents *out = reinterpret_cast<ClusterExtents *>(out_v);
    uint32_t s = 0;
    for (int64_t i = 0; i < ncl; i++) {
        out[i] = ClusterExtents{s, filt[i].count};
        s += filt[i].count;
    }
}

// Angle sort: src/main.cpp:737-750 (sort by (cluster index, slope); stable here).
void orc_angle_sort(void *pts_v, uint32_t *idx, int64_t n) {
    if (n <= 0) return;  // nothing to sort (and no memcpy from an empty vector's data())
    ClusterPoint *pts = reinterpret_cast<ClusterPoint *>(pts_v);
    std::vector<int64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    __gnu_parallel::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        if (idx[a] != idx[b]) return idx[a] < idx[b];
        return pts[a].slope < pts[b].slope;
    });
    std::vector<ClusterPoint> p2(n);
    std::vector<uint32_t> i2(n);
    for (int64_t i = 0; i < n; i++) { p2[i] = pts[order[i]]; i2[i] = idx[order[i]]; }
    std::memcpy(pts, p2.data(), sizeof(ClusterPoint) * n);
    std::memcpy(idx, i2.data(), sizeof(uint32_t) * n);
}

// Line-fit points: src/main.cpp:768-775
static void linefit_m(const void *pts_v, int64_t n, const uint8_t *grey, int64_t w, int64_t h,
                      void *out_v, bool fx) {
    const ClusterPoint *pts = reinterpret_cast<const ClusterPoint *>(pts_v);
    LineFitPoint *out = reinterpret_cast<LineFitPoint *>(out_v);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        out[i] = compute_initial_linefit(pts[i], (uint32_t)w, (uint32_t)h, grey, fx);
}
void orc_linefit(const void *pts_v, int64_t n, const uint8_t *grey, int64_t w, int64_t h,
                 void *out_v) {
    linefit_m(pts_v, n, grey, w, h, out_v, false);
}

// Moment prefix: src/main.cpp:793-801 (inclusive_scan_by_segment keyed by
// the cluster index; exact int64 here).
void orc_linefit_prefix(const uint32_t *idx, const void *in_v, int64_t n, void *out_v) {
    const LineFitPoint *in = reinterpret_cast<const LineFitPoint *>(in_v);
    LineFitPoint *out = reinterpret_cast<LineFitPoint *>(out_v);
    for (int64_t i = 0; i < n; i++) {
        if (i == 0 || idx[i] != idx[i - 1]) out[i] = in[i];
        else {
            LineFitPoint s = out[i - 1];
            s = lf_add(s, in[i]);
            s.a = s.a + in[i].a;
            s.b = s.b + in[i].b;
            out[i] = s;
        }
    }
}

// K10 fit_lines: src/fit_quad.cpp:153-375, emulated work-group by
// work-group (32 items, 24 new points each, three barrier phases). `out`
// (n entries) is zeroed here as src/main.cpp:316-317 does.
void orc_fit_lines(const void *pts_v, const uint32_t *cidx, const void *ext_v, int64_t n,
                   void *out_v) {
    const LineFitPoint *pts = reinterpret_cast<const LineFitPoint *>(pts_v);
    const ClusterExtents *ext = reinterpret_cast<const ClusterExtents *>(ext_v);
    Corner *out = reinterpret_cast<Corner *>(out_v);
    std::memset(out, 0, sizeof(Corner) * n);
    if (n <= 0) return;
    constexpr int64_t WG = 32, PPE = 3, END = PPE + 1;
    const int64_t groups = (n + 23) / 24;
    const float qnan = std::numeric_limits<float>::quiet_NaN();
#pragma omp parallel for
    for (int64_t g = 0; g < groups; g++) {
        uint32_t l_idx[WG];
        ClusterExtents l_ext[WG];
        float e_same[WG], e_diff[WG], f_same[WG], f_diff[WG];
        int64_t xs[WG];
        auto window_err = [&](const ClusterExtents &e, int64_t x) -> float {
            size_t k = std::min((size_t)20, (size_t)(e.count / 12));
            if (k < 2) return qnan;
            int64_t a = (int64_t)e.count + x - (int64_t)e.start - (int64_t)k;
            int64_t b = (int64_t)e.count + x - (int64_t)e.start + (int64_t)k;
            size_t i0 = (size_t)a % e.count, i1 = (size_t)b % e.count;
            size_t total;
            LineFitPoint m = get_moment(pts + e.start, e.count, i0, i1, &total,
                                        n - (int64_t)e.start, nullptr);
            float err;
            fit_line(m, total, nullptr, &err, nullptr);
            return err;
        };
        // phase 1: load indices and extents
        for (int64_t l = 0; l < WG; l++) {
            int64_t x = g * (WG - 2 * END) + l - PPE;
            xs[l] = x;
            int64_t cl = std::min(std::max(x, (int64_t)0), n - 1);
            l_idx[l] = cidx[cl];
            l_ext[l] = ext[l_idx[l]];
        }
        // phase 2: own-cluster and neighbour-cluster errors
        for (int64_t l = 0; l < WG; l++) {
            int64_t x = xs[l];
            e_same[l] = window_err(l_ext[l], x);
            int64_t their = l;
            if (l >= END && l_idx[l - END] != l_idx[l]) their = l - END;
            if (l < WG - END && l_idx[l + END] != l_idx[l]) their = l + END;
            e_diff[l] = (their != l) ? window_err(l_ext[their], x) : qnan;
        }
        // phase 3: 7-tap filter
        for (int64_t l = PPE; l < WG - PPE; l++) {
            float acc = 0.0f, dacc = 0.0f;
            for (int i = 0; i < 7; i++) {
                int64_t c = l + i - PPE;
                if (l_idx[c] == l_idx[l]) {
                    float t1 = e_same[c] * FILTER_DATA[i];
                    acc = acc + t1;
                    float t2 = e_diff[c] * FILTER_DATA[i];
                    dacc = dacc + t2;
                } else {
                    float t1 = e_same[c] * FILTER_DATA[i];
                    dacc = dacc + t1;
                    float t2 = e_diff[c] * FILTER_DATA[i];
                    acc = acc + t2;
                }
            }
            f_same[l] = acc;
            f_diff[l] = dacc;
        }
        // phase 4: strict local maxima
        for (int64_t l = END; l < WG - END; l++) {
            int64_t x = xs[l];
            if (x < 0 || x >= n) continue;
            float cur = f_same[l];
            bool before_same = l_idx[l] == l_idx[l - 1];
            bool gb = cur > (before_same ? f_same : f_diff)[l - 1];
            bool after_same = l_idx[l] == l_idx[l + 1];
            bool ga = cur > (after_same ? f_same : f_diff)[l + 1];
            if (gb && ga)
                out[x] = Corner{(uint32_t)(x - (int64_t)l_ext[l].start), l_idx[l], cur};
        }
    }
}

// Corner segments: src/main.cpp:832-907. Returns the segment count;
// *n_corner receives the compacted corner count.
int64_t orc_corner_segments(const void *corners_v, int64_t n, void *compacted_v, void *seg_v,
                            int64_t *n_corner) {
    const Corner *c = reinterpret_cast<const Corner *>(corners_v);
    Corner *cc = reinterpret_cast<Corner *>(compacted_v);
    PeakExtents *seg = reinterpret_cast<PeakExtents *>(seg_v);
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (c[i].error != 0) cc[m++] = c[i];
    *n_corner = m;
    std::stable_sort(cc, cc + m, [](const Corner &a, const Corner &b) {
        if (a.cluster_index != b.cluster_index) return a.cluster_index < b.cluster_index;
        return a.error > b.error;
    });
    int64_t e = 0;
    for (int64_t i = 0; i < m; i++) {
        if (i == 0 || cc[i].cluster_index != cc[i - 1].cluster_index) seg[e++] = PeakExtents{(uint32_t)i, 1};
        else seg[e - 1].count += 1;
    }
    return e;
}

// src/fit_quad.cpp:377-389
static void sort_4(size_t a[4]) {
    auto sw = [&](int i, int j) { if (!(a[i] < a[j])) std::swap(a[i], a[j]); };
    sw(0, 1); sw(2, 3); sw(0, 2); sw(1, 3); sw(1, 2);
}

// K11 do_indexing: src/fit_quad.cpp:392-535. One emulated work-group of 256
// combination numbers per peak segment; segment e is paired with
// cluster_extents[e], as the reference does. `fitted` (E entries) is zeroed here as
// src/main.cpp:318-319 does. n_points bounds reads of `points`.
static int64_t do_indexing_m(const void *ext_v, int64_t E, const void *corners_v,
                             const void *pts_v, int64_t n_points, const void *cext_v,
                             void *fitted_v, bool fx) {
    const PeakExtents *ext = reinterpret_cast<const PeakExtents *>(ext_v);
    const Corner *corners = reinterpret_cast<const Corner *>(corners_v);
    const LineFitPoint *pts = reinterpret_cast<const LineFitPoint *>(pts_v);
    const ClusterExtents *cext = reinterpret_cast<const ClusterExtents *>(cext_v);
    FittedQuad *fitted = reinterpret_cast<FittedQuad *>(fitted_v);
    std::memset(fitted, 0, sizeof(FittedQuad) * E);
    const float cos_crit = (float)std::cos(10 * M_PI / 180);
    const float MAX_MSE = 10.0f;
    const Combinadics &C = combinadics();
    uint64_t oob_total = 0;
#pragma omp parallel for reduction(+ : oob_total)
    for (int64_t e = 0; e < E; e++) {
        PeakExtents pk = ext[e];
        // fixed mode: the segment's own cluster
        ClusterExtents ce = fx ? cext[corners[pk.start].cluster_index] : cext[e];
        float cur_err[256];
        FittedQuad cand[256];
        uint64_t oob = 0;
        const int64_t bound = n_points - (int64_t)ce.start;
        size_t ncomb = C.n_choose_k(std::min(pk.count, 10u), 4);
        for (size_t comb = 0; comb < 256; comb++) {
            bool valid = comb < ncomb;
            float total_err = 0.0f;
            bool valid_fit = valid;
            FittedQuad q{};
            if (valid) {
                size_t ind[4];
                C.decode((uint8_t)comb, ind);
                size_t li[4];
                for (int k = 0; k < 4; k++) li[k] = corners[pk.start + ind[k]].line_fit_point_index;
                sort_4(li);
                float p01[4], p12[4];
                for (int s = 0; s < 4; s++) {
                    size_t total;
                    LineFitPoint m = get_moment(pts + ce.start, ce.count, li[s], li[(s + 1) & 3],
                                                &total, bound, &oob);
                    float err, mse;
                    fit_line(m, total, s == 0 ? p01 : (s == 1 ? p12 : nullptr), &err, &mse);
                    valid_fit &= mse <= MAX_MSE;
                    q.moments[s] = m;
                    q.num_in_moments[s] = (uint16_t)total;
                    total_err = total_err + err;
                }
                for (int k = 0; k < 4; k++) q.indices[k] = (uint16_t)li[k];
                float d1 = p01[2] * p12[2], d2 = p01[3] * p12[3];
                float dot = d1 + d2;
                valid_fit &= std::fabs(dot) <= cos_crit;
            }
            cur_err[comb] = valid_fit ? total_err : std::numeric_limits<float>::infinity();
            cand[comb] = q;
        }
        oob_total += oob;
        float lowest = std::numeric_limits<float>::infinity();
        for (int i = 0; i < 256; i++) lowest = std::min(lowest, cur_err[i]);
        if (lowest == std::numeric_limits<float>::infinity()) continue;
        for (int i = 0; i < 256; i++) {
            float d = cur_err[i] - lowest;
            if ((double)d < 0.1) {
                fitted[e] = cand[i];
                break;
            }
        }
    }
    return (int64_t)oob_total;
}
int64_t orc_do_indexing(const void *ext_v, int64_t E, const void *corners_v, const void *pts_v,
                        int64_t n_points, const void *cext_v, void *fitted_v) {
    return do_indexing_m(ext_v, E, corners_v, pts_v, n_points, cext_v, fitted_v, false);
}

// process_fitted_quads: src/main.cpp:1035-1162, minus decode_quads.
// Quads whose indices are all zero (never written) are skipped
// (src/main.cpp:996-1018, 1044-1047). Returns the number written.
int64_t orc_process_quads(const void *fq_v, int64_t E, void *out_v) {
    const FittedQuad *fq = reinterpret_cast<const FittedQuad *>(fq_v);
    QuadCorners *out = reinterpret_cast<QuadCorners *>(out_v);
    const int min_tag_width = 8;
    const double cos_crit = std::cos(10 * M_PI / 180);
    int64_t m = 0;
    for (int64_t qi = 0; qi < E; qi++) {
        const FittedQuad &q = fq[qi];
        if (!(q.indices[0] != 0 || q.indices[1] != 0 || q.indices[2] != 0 || q.indices[3] != 0))
            continue;
        QuadCorners c{};
        c.reversed_border = 0;
        c.blob_index = (uint32_t)qi;
        float lines[4][4];
        for (int i = 0; i < 4; i++) {
            float err, mse;
            fit_line(q.moments[i], q.num_in_moments[i], lines[i], &err, &mse);
        }
        bool bad = false;
        for (int i = 0; i < 4; i++) {
            double A00 = lines[i][3], A01 = -lines[(i + 1) & 3][3];
            double A10 = -lines[i][2], A11 = lines[(i + 1) & 3][2];
            double B0 = -lines[i][0] + lines[(i + 1) & 3][0];
            double B1 = -lines[i][1] + lines[(i + 1) & 3][1];
            double t1 = A00 * A11, t2 = A10 * A01;
            double det = t1 - t2;
            double W00 = A11 / det, W01 = -A01 / det;
            if (std::fabs(det) < 0.001) { bad = true; break; }
            double u1 = W00 * B0, u2 = W01 * B1;
            double L0 = u1 + u2;
            double v0 = L0 * A00, v1 = L0 * A10;
            c.corners[i][0] = (float)((double)lines[i][0] + v0);
            c.corners[i][1] = (float)((double)lines[i][1] + v1);
        }
        if (bad) continue;
        {
            float area = 0;
            float length[3], p;
            for (int i = 0; i < 3; i++) {
                int a = i, b = (i + 1) % 3;
                length[i] = hypot_f(c.corners[b][0] - c.corners[a][0], c.corners[b][1] - c.corners[a][1]);
            }
            p = (length[0] + length[1] + length[2]) / 2;
            {
                float t = p * (p - length[0]);
                t = t * (p - length[1]);
                t = t * (p - length[2]);
                area += std::sqrt(t);
            }
            const int idxs[] = {2, 3, 0, 2};
            for (int i = 0; i < 3; i++) {
                int a = idxs[i], b = idxs[i + 1];
                length[i] = hypot_f(c.corners[b][0] - c.corners[a][0], c.corners[b][1] - c.corners[a][1]);
            }
            p = (length[0] + length[1] + length[2]) / 2;
            {
                float t = p * (p - length[0]);
                t = t * (p - length[1]);
                t = t * (p - length[2]);
                area += std::sqrt(t);
            }
            if ((double)area < 0.95 * min_tag_width * min_tag_width) continue;
        }
        {
            bool reject = false;
            for (int i = 0; i < 4; i++) {
                int i0 = i, i1 = (i + 1) & 3, i2 = (i + 2) & 3;
                float dx1 = c.corners[i1][0] - c.corners[i0][0];
                float dy1 = c.corners[i1][1] - c.corners[i0][1];
                float dx2 = c.corners[i2][0] - c.corners[i1][0];
                float dy2 = c.corners[i2][1] - c.corners[i1][1];
                float n1 = dx1 * dx2, n2 = dy1 * dy2;
                float num = n1 + n2;
                float s1a = dx1 * dx1, s1b = dy1 * dy1, s2a = dx2 * dx2, s2b = dy2 * dy2;
                float s1 = s1a + s1b, s2 = s2a + s2b;
                float den = std::sqrt(s1 * s2);
                float cosd = num / den;
                float c1 = dx1 * dy2, c2 = dy1 * dx2;
                if ((double)std::fabs(cosd) > cos_crit || c1 < c2) { reject = true; break; }
            }
            if (reject) continue;
        }
        out[m++] = c;
    }
    return m;
}

// ======================================================================
// Fixed mode. The functions above restate the reference; these correct
// seven of its behaviours and have no reference counterpart, so they are
// pinned against textbook definitions instead (tests/test_oracle_fixed.py):
//  * CCL: pixel-level, white 8-connected, black 4-connected, 127 pixels
//    unlabelled. The reference's block graph (src/segmentation.cpp:139-255)
//    links colour sets through a colour-agnostic neighbour mask.
//  * labels: white 0x80000000 | (root + 1), black root + 1, 127 -> 0, with
//    root the component's minimum pixel index; `sizes` has w*h + 1 entries
//    and sizes[0] = 0, so neither 127 pixels nor out-of-image neighbours
//    (both read as label 0) are ever a large component. The reference
//    counts both into sizes[0] (src/find_boundaries.cpp:49-55, 82).
//  * boundaries: column 0 is a local pixel like every other (the reference's
//    tiling skips it, src/find_boundaries.cpp:66-69); records are kept by
//    key, so the record of pixel (0, 0) RIGHT white-to-black, whose 32 bits
//    are zero, is kept.
//  * records decode both half-pixel bits: x2 = 2x + dx, y2 = 2y + dy,
//    (gx, gy) = +-(dx, dy) (include/boundary_point.h:16 masks with &).
//  * the line-fit weight samples pixel (rx/2, ry/2) (include/line_fit_point.h:63
//    takes the row from rx).
//  * fit_lines: global index 0 is a corner candidate like every other point
//    (src/fit_quad.cpp:184-187 never evaluates it).
//  * do_indexing: a peak segment is fitted against its own cluster's
//    extents (src/fit_quad.cpp:405-408 takes cluster_extents[e]).

// Root of n (the minimum of its class: parents are always smaller), halving
// the path; the result does not depend on the tree's shape.
static uint32_t fx_find(uint32_t *par, uint32_t n) {
    while (par[n] != n) {
        par[n] = par[par[n]];
        n = par[n];
    }
    return n;
}
static void fx_union(uint32_t *par, uint32_t a, uint32_t b) {
    a = fx_find(par, a);
    b = fx_find(par, b);
    if (a < b) par[b] = a;
    else if (b < a) par[a] = b;
}

// thr (w*h, values 0/127/255) -> labels (w*h), sizes (w*h + 1).
void orc_fx_segmentation(const uint8_t *thr, int64_t w, int64_t h, uint32_t *labels,
                         uint32_t *sizes) {
    const int64_t n = w * h;
    std::vector<uint32_t> par((size_t)n);
    std::iota(par.begin(), par.end(), 0u);
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) {
            const int64_t p = y * w + x;
            const uint8_t v = thr[p];
            if (v == 127) continue;
            auto link = [&](int64_t q) {
                if (thr[q] == v) fx_union(par.data(), (uint32_t)p, (uint32_t)q);
            };
            if (x > 0) link(p - 1);
            if (y > 0) {
                link(p - w);
                if (v == 255) {
                    if (x > 0) link(p - w - 1);
                    if (x + 1 < w) link(p - w + 1);
                }
            }
        }
    std::memset(sizes, 0, sizeof(uint32_t) * (size_t)(n + 1));
    for (int64_t p = 0; p < n; p++) {
        const uint8_t v = thr[p];
        if (v == 127) {
            labels[p] = 0;
            continue;
        }
        const uint32_t r = fx_find(par.data(), (uint32_t)p);
        labels[p] = (v == 255 ? LABEL_PIXEL_MASK : 0u) | (r + 1);
        sizes[r + 1]++;
    }
}

// The reference's four record planes (RIGHT, BOTTOM_LEFT with its duplicate
// rule, BOTTOM, BOTTOM_RIGHT; src/find_boundaries.cpp:94-170) for every
// pixel, column 0 included. points/keys: 4*w*h, pre-filled like the
// reference's (0 / all ones).
void orc_fx_find_boundaries(const uint32_t *labels, const uint32_t *sizes, int64_t w, int64_t h,
                            uint32_t *points_out, uint64_t *keys) {
    BoundaryPoint *points = reinterpret_cast<BoundaryPoint *>(points_out);
    const int64_t n = w * h;
    std::memset(points, 0, sizeof(BoundaryPoint) * 4 * n);
    std::memset(keys, 0xFF, sizeof(uint64_t) * 4 * n);
    auto at = [&](int64_t x, int64_t y) -> uint32_t {
        return (x < 0 || x >= w || y < 0 || y >= h) ? 0u : labels[y * w + x];
    };
    auto big = [&](uint32_t l) { return sizes[l & LABEL_VALUE_MASK] >= 25; };
#pragma omp parallel for
    for (int64_t y = 0; y < h; y++)
        for (int64_t x = 0; x < w; x++) {
            const uint32_t L = at(x, y);
            const uint32_t lv = L & LABEL_VALUE_MASK, wl = L & LABEL_PIXEL_MASK;
            auto handle = [&](int dy, int dx, int64_t plane, uint16_t hp) {
                const uint32_t T = at(x + dx, y + dy);
                const uint32_t wt = T & LABEL_PIXEL_MASK, tv = T & LABEL_VALUE_MASK;
                if (wt == wl || !big(L) || !big(T)) return;
                const int64_t oi = n * plane + y * w + x;
                points[oi] = BoundaryPoint{(uint16_t)((uint16_t)x | (uint16_t)(hp << 14)),
                                           wt > wl ? (uint16_t)(COORDINATE_COLOR_DIRECTION_MASK | y)
                                                   : (uint16_t)y};
                keys[oi] = ((uint64_t)std::min(lv, tv) << 32) | (uint64_t)std::max(lv, tv);
            };
            handle(0, 1, 0, RIGHT);
            const uint32_t left = at(x - 1, y), down = at(x, y + 1);
            if (!((left & LABEL_PIXEL_MASK) != (down & LABEL_PIXEL_MASK) && big(left) && big(down)))
                handle(1, -1, 1, BOTTOM_LEFT);
            handle(1, 0, 2, BOTTOM);
            handle(1, 1, 3, BOTTOM_RIGHT);
        }
}

// Compaction on the key (every emitted record has a key below all ones)
// and the stable key sort. Returns N_bp.
int64_t orc_fx_compact_sort(const uint32_t *points_sparse, const uint64_t *keys_sparse,
                            int64_t n_sparse, uint32_t *out_points, uint64_t *out_keys) {
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n_sparse; i++)
        if (keys_sparse[i] != ~0ull) idx.push_back(i);
    __gnu_parallel::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        return keys_sparse[a] < keys_sparse[b];
    });
    for (size_t i = 0; i < idx.size(); i++) {
        out_points[i] = points_sparse[idx[i]];
        out_keys[i] = keys_sparse[idx[i]];
    }
    return (int64_t)idx.size();
}

int64_t orc_fx_cluster_bounds(const uint32_t *pts_u32, const uint64_t *keys, int64_t n,
                              void *out_v) {
    return cluster_bounds_m(pts_u32, keys, n, out_v, true);
}
int64_t orc_fx_cluster_points(const uint32_t *pts_u32, int64_t n, const void *filt_v,
                              int64_t ncl, void *out_pts_v, uint32_t *out_idx) {
    return cluster_points_m(pts_u32, n, filt_v, ncl, out_pts_v, out_idx, true);
}
void orc_fx_linefit(const void *pts_v, int64_t n, const uint8_t *grey, int64_t w, int64_t h,
                    void *out_v) {
    linefit_m(pts_v, n, grey, w, h, out_v, true);
}

// fit_lines as the per-cluster cyclic computation, for every point: window
// errors e(p) of the 2k+1 points centred on p (k = min(20, n/12), none when
// k < 2), f(p) = sum_t F[t] e(p+t-3) accumulated in order t = 0..6, and a
// corner where f(p) > f(p-1) and f(p) > f(p+1), indices modulo the cluster
// size n. `out` (n entries) is zeroed first.
void orc_fx_fit_lines(const void *pts_v, const uint32_t *cidx, const void *ext_v, int64_t n,
                      void *out_v) {
    const LineFitPoint *pts = reinterpret_cast<const LineFitPoint *>(pts_v);
    const ClusterExtents *ext = reinterpret_cast<const ClusterExtents *>(ext_v);
    Corner *out = reinterpret_cast<Corner *>(out_v);
    std::memset(out, 0, sizeof(Corner) * n);
#pragma omp parallel for
    for (int64_t x = 0; x < n; x++) {
        const ClusterExtents e = ext[cidx[x]];
        const size_t cnt = e.count, k = std::min((size_t)20, cnt / 12);
        if (k < 2) continue;
        const size_t j = (size_t)x - e.start;
        auto err_at = [&](size_t pos) {
            size_t total;
            LineFitPoint m = get_moment(pts + e.start, cnt, (pos + cnt - k) % cnt, (pos + k) % cnt,
                                        &total, (int64_t)cnt, nullptr);
            float err;
            fit_line(m, total, nullptr, &err, nullptr);
            return err;
        };
        float er[9];
        for (int d = 0; d < 9; d++) er[d] = err_at((j + 4 * cnt + d - 4) % cnt);
        float f[3];
        for (int q = 0; q < 3; q++) {
            float acc = 0.0f;
            for (int t = 0; t < 7; t++) {
                float t1 = er[q + t] * FILTER_DATA[t];
                acc = acc + t1;
            }
            f[q] = acc;
        }
        if (f[1] > f[0] && f[1] > f[2]) out[x] = Corner{(uint32_t)j, cidx[x], f[1]};
    }
}

int64_t orc_fx_do_indexing(const void *ext_v, int64_t E, const void *corners_v, const void *pts_v,
                           int64_t n_points, const void *cext_v, void *fitted_v) {
    return do_indexing_m(ext_v, E, corners_v, pts_v, n_points, cext_v, fitted_v, true);
}

// Whole front end (src/main.cpp:322-1023 minus debug dumps and decoding) on
// one grey frame, decimated by f first. `out` receives up to max_out quads;
// the return value is the total number of quads, or -1 when the decimated
// frame is below 8x8 or not a multiple of 4 in each dimension: the
// threshold's last pass indexes the filtered tile extrema as
// (y / 4) * (w / 4) + x / 4 (src/threshold.cpp:71-78), which reads past them
// for such dims, so the library rejects them too (ATQ_ERR_BAD_DIMS).
static int64_t detect_m(const uint8_t *gray_in, int64_t w_in, int64_t h_in, int f, int nthreads,
                        void *out, int64_t max_out, orc_stats *st, bool fx) {
    {
        const int64_t wd = f > 1 ? 1 + (w_in - 1) / f : w_in, hd = f > 1 ? 1 + (h_in - 1) / f : h_in;
        if (w_in < 1 || h_in < 1 || wd < 8 || hd < 8 || (wd & 3) || (hd & 3)) return -1;
    }
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    omp_set_num_threads(nthreads_or_default(nthreads));
#else
    (void)nthreads;
#endif
    std::vector<uint8_t> dec;
    const uint8_t *g = gray_in;
    int64_t w = w_in, h = h_in;
    if (f > 1) {
        w = 1 + (w_in - 1) / f;
        h = 1 + (h_in - 1) / f;
        dec.resize((size_t)(w * h));
        orc_decimate(gray_in, w_in, h_in, f, dec.data());
        g = dec.data();
    }
    const int64_t n = w * h;
    std::vector<uint8_t> ext((size_t)(2 * (w / 4) * (h / 4))), fil(ext.size()), thr((size_t)n);
    orc_threshold(g, w, h, ext.data(), fil.data(), thr.data());
    std::vector<uint32_t> scratch((size_t)n), labels((size_t)n), sizes((size_t)n + 1);
    if (fx) orc_fx_segmentation(thr.data(), w, h, labels.data(), sizes.data());
    else orc_segmentation(thr.data(), w, h, scratch.data(), labels.data(), sizes.data());
    std::vector<uint32_t> sp((size_t)(4 * n));
    std::vector<uint64_t> sk((size_t)(4 * n));
    if (fx) orc_fx_find_boundaries(labels.data(), sizes.data(), w, h, sp.data(), sk.data());
    else orc_find_boundaries(labels.data(), sizes.data(), w, h, sp.data(), sk.data());
    std::vector<uint32_t> cp((size_t)(4 * n));
    std::vector<uint64_t> ck((size_t)(4 * n));
    int64_t nbp = fx ? orc_fx_compact_sort(sp.data(), sk.data(), 4 * n, cp.data(), ck.data())
                     : orc_compact_sort(sp.data(), sk.data(), 4 * n, cp.data(), ck.data());
    sp.clear(); sp.shrink_to_fit(); sk.clear(); sk.shrink_to_fit();
    std::vector<ClusterBounds> bounds((size_t)std::max<int64_t>(nbp, 1)), filt(bounds.size());
    int64_t npairs = cluster_bounds_m(cp.data(), ck.data(), nbp, bounds.data(), fx);
    int64_t ncl = orc_filter_bounds(bounds.data(), npairs, w, h, filt.data());
    std::vector<ClusterPoint> fpts((size_t)std::max<int64_t>(nbp, 1));
    std::vector<uint32_t> fidx(fpts.size());
    int64_t nfp = cluster_points_m(cp.data(), nbp, filt.data(), ncl, fpts.data(), fidx.data(), fx);
    std::vector<ClusterExtents> cext((size_t)std::max<int64_t>(ncl, 1));
    orc_cluster_extents(filt.data(), ncl, cext.data());
    orc_angle_sort(fpts.data(), fidx.data(), nfp);
    std::vector<LineFitPoint> lf((size_t)std::max<int64_t>(nfp, 1)), lfp(lf.size());
    linefit_m(fpts.data(), nfp, g, w, h, lf.data(), fx);
    orc_linefit_prefix(fidx.data(), lf.data(), nfp, lfp.data());
    std::vector<Corner> corners((size_t)std::max<int64_t>(nfp, 1)), cc(corners.size());
    if (fx) orc_fx_fit_lines(lfp.data(), fidx.data(), cext.data(), nfp, corners.data());
    else orc_fit_lines(lfp.data(), fidx.data(), cext.data(), nfp, corners.data());
    std::vector<PeakExtents> seg(corners.size());
    int64_t ncorner = 0;
    int64_t E = orc_corner_segments(corners.data(), nfp, cc.data(), seg.data(), &ncorner);
    std::vector<FittedQuad> fq((size_t)std::max<int64_t>(E, 1));
    int64_t oob = do_indexing_m(seg.data(), E, cc.data(), lfp.data(), nfp, cext.data(), fq.data(), fx);
    std::vector<QuadCorners> quads((size_t)std::max<int64_t>(E, 1));
    int64_t nq = orc_process_quads(fq.data(), E, quads.data());
    if (out && max_out > 0 && nq > 0)
        std::memcpy(out, quads.data(), sizeof(QuadCorners) * (size_t)std::min(nq, max_out));
    if (st) {
        st->width = w; st->height = h;
        st->n_sparse_points = nbp; st->n_pairs = npairs; st->n_clusters = ncl;
        st->n_cluster_points = nfp; st->n_corners = ncorner; st->n_segments = E;
        st->n_quads = nq; st->oob_reads = oob;
        int64_t nf = 0;
        for (int64_t i = 0; i < E; i++)
            if (fq[i].indices[0] | fq[i].indices[1] | fq[i].indices[2] | fq[i].indices[3]) nf++;
        st->n_fitted = nf;
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return nq;
}
int64_t orc_detect(const uint8_t *gray_in, int64_t w_in, int64_t h_in, int f, int nthreads,
                   void *out, int64_t max_out, orc_stats *st) {
    return detect_m(gray_in, w_in, h_in, f, nthreads, out, max_out, st, false);
}
// The same in fixed mode.
int64_t orc_fx_detect(const uint8_t *gray_in, int64_t w_in, int64_t h_in, int f, int nthreads,
                      void *out, int64_t max_out, orc_stats *st) {
    return detect_m(gray_in, w_in, h_in, f, nthreads, out, max_out, st, true);
}

// Single-call helpers used by tests.
void orc_fit_line(const void *m, int64_t num, float *line, float *err, float *mse) {
    fit_line(*reinterpret_cast<const LineFitPoint *>(m), (size_t)num, line, err, mse);
}
void orc_get_moment(const void *pts, int64_t n, int64_t i0, int64_t i1, void *out, int64_t *total) {
    size_t t;
    *reinterpret_cast<LineFitPoint *>(out) =
        get_moment(reinterpret_cast<const LineFitPoint *>(pts), (size_t)n, (size_t)i0, (size_t)i1,
                   &t, n, nullptr);
    *total = (int64_t)t;
}
float orc_calc_theta(uint16_t px, uint16_t py, float cx, float cy) {
    return BoundaryPoint{px, py}.calc_theta(cx, cy);
}
float orc_fx_calc_theta(uint16_t px, uint16_t py, float cx, float cy) {
    return BoundaryPoint{px, py}.calc_theta(cx, cy, true);
}
void orc_fx_bounds_from_point(uint16_t px, uint16_t py, uint32_t start, void *out) {
    *reinterpret_cast<ClusterBounds *>(out) = ClusterBounds::from_point(BoundaryPoint{px, py}, start, true);
}
void orc_fx_initial_linefit(uint16_t px, uint16_t py, int64_t w, int64_t h, const uint8_t *grey,
                            void *out) {
    *reinterpret_cast<LineFitPoint *>(out) =
        compute_initial_linefit(ClusterPoint{px, py, 0.0f}, (uint32_t)w, (uint32_t)h, grey, true);
}
void orc_bounds_cx_cy_dot(const void *b, float *out3) {
    const ClusterBounds *c = reinterpret_cast<const ClusterBounds *>(b);
    out3[0] = c->cx(); out3[1] = c->cy(); out3[2] = c->dot();
}
int orc_valid_blob(const void *b, int64_t w, int64_t h) {
    return ValidBlobFilter((size_t)w, (size_t)h)(*reinterpret_cast<const ClusterBounds *>(b)) ? 1 : 0;
}
void orc_bounds_from_point(uint16_t px, uint16_t py, uint32_t start, void *out) {
    *reinterpret_cast<ClusterBounds *>(out) = ClusterBounds::from_point(BoundaryPoint{px, py}, start);
}
void orc_reduce_bounds(const void *a, const void *b, void *out) {
    *reinterpret_cast<ClusterBounds *>(out) = reduce_bounds(*reinterpret_cast<const ClusterBounds *>(a),
                                                            *reinterpret_cast<const ClusterBounds *>(b));
}
void orc_initial_linefit(uint16_t px, uint16_t py, int64_t w, int64_t h, const uint8_t *grey, void *out) {
    *reinterpret_cast<LineFitPoint *>(out) =
        compute_initial_linefit(ClusterPoint{px, py, 0.0f}, (uint32_t)w, (uint32_t)h, grey);
}
int orc_combinadics_n_choose_k(int n, int k) { return combinadics().n_choose_k((size_t)n, (size_t)k); }
void orc_combinadics_decode(int num, int64_t *out4) {
    size_t o[4];
    combinadics().decode((uint8_t)num, o);
    for (int i = 0; i < 4; i++) out4[i] = (int64_t)o[i];
}
float orc_hypotf(float a, float b) { return hypot_f(a, b); }
int orc_max_threads(void) { return nthreads_or_default(0); }

}  // extern "C"
