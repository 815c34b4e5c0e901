// decode.hip - tag payloads of quads (the table-free part of the apriltag
// library's quad_decode, which the reference runs per quad through
// quad_decode_task, src/quad_decoding.cpp:46-75).
//
// One 64-lane work-group per quad:
//  * the homography H from the tag square [-1,1]^2 to the quad (corner i at
//    (-1,-1), (1,-1), (1,1), (-1,1)), in closed form: the projective map of
//    the unit square onto the quad (Heckbert's square-to-quad mapping)
//    composed with [-1,1] -> [0,1];
//  * 8 lines of width_at_border border samples (just outside: white, just
//    inside: black), each pixel read at the truncated projected position,
//    one sample per lane; lane 0 sums each model's normal equations in
//    sample order and solves the 3x3 system (Cholesky) for the plane
//    I = c0 x + c1 y + c2;
//  * the polarity test: white - black at the centre must be >= 0 (< 0 for a
//    reversed border); a model whose samples span no plane (too few inside
//    the image) fails it too;
//  * one bilinear sample per bit cell (pixel centres at +0.5; cells whose
//    four pixels are not all inside stay 0) minus the mean of the two
//    planes, as float in a total_width^2 grid, then v += s * (4 v - the four
//    neighbours) with zero outside the grid;
//  * lane 0 reads the bits in layout order (1 where v > 0) and the margin
//    min(white sum / (1 + white bits), black sum / (1 + black bits)).
// Matching: one thread per code, the four rotated copies of the code (bit
// permutations from the layout) against every table entry.
#include "kernels.hpp"

namespace atq {

namespace {

struct DecodeLayout {
    int32_t wb, tw, rev, nbits;
    float sharp;
    uint8_t bx[ATQ_TAG_MAX_BITS], by[ATQ_TAG_MAX_BITS];
};

struct Homography {
    double m[9];
    ATQ_DEV void project(double x, double y, double &px, double &py) const {
        const double d = m[6] * x + m[7] * y + m[8];
        px = (m[0] * x + m[1] * y + m[2]) / d;
        py = (m[3] * x + m[4] * y + m[5]) / d;
    }
};

// Square-to-quad map: (u, v) in [0,1]^2, (0,0) -> p0, (1,0) -> p1, (1,1) -> p2,
// (0,1) -> p3, then u = (x + 1) / 2, v = (y + 1) / 2, scaled to m[8] = 1.
ATQ_DEV Homography quad_homography(const float (&c)[4][2]) {
    const double x0 = c[0][0], y0 = c[0][1], x1 = c[1][0], y1 = c[1][1];
    const double x2 = c[2][0], y2 = c[2][1], x3 = c[3][0], y3 = c[3][1];
    const double sx = x0 - x1 + x2 - x3, sy = y0 - y1 + y2 - y3;
    double a, b, cc, d, e, f, g, h;
    if (sx == 0.0 && sy == 0.0) {  // a parallelogram: affine
        a = x1 - x0;
        b = x3 - x0;
        d = y1 - y0;
        e = y3 - y0;
        g = h = 0.0;
    } else {
        const double dx1 = x1 - x2, dx2 = x3 - x2, dy1 = y1 - y2, dy2 = y3 - y2;
        const double den = dx1 * dy2 - dx2 * dy1;
        g = (sx * dy2 - dx2 * sy) / den;
        h = (dx1 * sy - sx * dy1) / den;
        a = x1 - x0 + g * x1;
        b = x3 - x0 + h * x3;
        d = y1 - y0 + g * y1;
        e = y3 - y0 + h * y3;
    }
    cc = x0;
    f = y0;
    // [a b c; d e f; g h 1] * [0.5 0 0.5; 0 0.5 0.5; 0 0 1]
    Homography H;
    const double k = 0.5 * g + 0.5 * h + 1.0;
    H.m[0] = 0.5 * a / k;
    H.m[1] = 0.5 * b / k;
    H.m[2] = (0.5 * a + 0.5 * b + cc) / k;
    H.m[3] = 0.5 * d / k;
    H.m[4] = 0.5 * e / k;
    H.m[5] = (0.5 * d + 0.5 * e + f) / k;
    H.m[6] = 0.5 * g / k;
    H.m[7] = 0.5 * h / k;
    H.m[8] = 1.0;
    return H;
}

// Plane c0 x + c1 y + c2 through the normal equations (upper triangle a:
// xx, xy, x, yy, y, n; b: xv, yv, v) by Cholesky; false when the samples do
// not span a plane (a pivot at or below 1e-9).
ATQ_DEV bool solve_plane(const double (&a)[6], const double (&b)[3], double (&c)[3]) {
    const double p0 = a[0];
    if (!(p0 > 1e-9)) return false;
    const double l00 = sqrt(p0);
    const double l10 = a[1] / l00, l20 = a[2] / l00;
    const double p1 = a[3] - l10 * l10;
    if (!(p1 > 1e-9)) return false;
    const double l11 = sqrt(p1);
    const double l21 = (a[4] - l20 * l10) / l11;
    const double p2 = a[5] - l20 * l20 - l21 * l21;
    if (!(p2 > 1e-9)) return false;
    const double l22 = sqrt(p2);
    const double z0 = b[0] / l00;
    const double z1 = (b[1] - l10 * z0) / l11;
    const double z2 = (b[2] - l20 * z0 - l21 * z1) / l22;
    c[2] = z2 / l22;
    c[1] = (z1 - l21 * c[2]) / l11;
    c[0] = (z0 - l10 * c[1] - l20 * c[2]) / l00;
    return true;
}

ATQ_DEV double plane_at(const double (&c)[3], double x, double y) { return c[0] * x + c[1] * y + c[2]; }

ATQ_DEV double bilinear(const uint8_t *gray, uint32_t w, uint32_t h, double px, double py) {
    const double fx0 = floor(px - 0.5), fy0 = floor(py - 0.5);
    const double cx0 = ceil(px - 0.5), cy0 = ceil(py - 0.5);
    if (!(fx0 >= 0.0 && fy0 >= 0.0 && cx0 < (double)w && cy0 < (double)h)) return -1.0;  // NaN too
    const uint32_t x1 = (uint32_t)fx0, x2 = (uint32_t)cx0, y1 = (uint32_t)fy0, y2 = (uint32_t)cy0;
    const double x = px - 0.5 - fx0, y = py - 0.5 - fy0;
    return gray[(size_t)y1 * w + x1] * (1 - x) * (1 - y) + gray[(size_t)y1 * w + x2] * x * (1 - y) +
           gray[(size_t)y2 * w + x1] * (1 - x) * y + gray[(size_t)y2 * w + x2] * x * y;
}

constexpr int MAX_SAMPLES = 8 * ATQ_TAG_MAX_WIDTH;
constexpr int MAX_CELLS = ATQ_TAG_MAX_WIDTH * ATQ_TAG_MAX_WIDTH;

__global__ void __launch_bounds__(64)
k_decode(const uint8_t *__restrict__ gray, uint32_t w, uint32_t h,
         const atq_quad_corners *__restrict__ quads, uint32_t n_cap, const uint32_t *n_dev,
         DecodeLayout L, atq_tag_code *__restrict__ out) {
    __shared__ double sx[MAX_SAMPLES], sy[MAX_SAMPLES], sv[MAX_SAMPLES];
    __shared__ int skind[MAX_SAMPLES];  // 0 outside the image, 1 white, 2 black
    __shared__ float vals[MAX_CELLS], sharp[MAX_CELLS];
    __shared__ double model[6];         // white c0..c2, black c0..c2
    __shared__ int polarity_ok;
    const uint32_t n = n_dev ? min(*n_dev, n_cap) : n_cap;
    const int lane = threadIdx.x;
    const int wb = L.wb, tw = L.tw, ns = 8 * wb, mn = (wb - tw) / 2;
    for (uint32_t q = blockIdx.x; q < n; q += gridDim.x) {
        float c[4][2];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            c[i][0] = quads[q].corners[i][0];
            c[i][1] = quads[q].corners[i][1];
        }
        const Homography H = quad_homography(c);
        // border samples: pattern p (white outside / black inside the left,
        // right, top and bottom edges), sample i along it
        for (int s = lane; s < ns; s += 64) {
            // p: left white, left black, right white, right black, then the
            // same for the top and bottom rows (white half a cell outside
            // the border, black half a cell inside)
            const int p = s / wb, i = s % wb;
            const bool vertical = p < 4, white = (p & 1) == 0, far = (p & 2) != 0;
            const double across = far ? (white ? wb + 0.5 : wb - 0.5) : (white ? -0.5 : 0.5);
            const double along = 0.5 + i;
            const double bxv = vertical ? across : along, byv = vertical ? along : across;
            const double tx = 2.0 * (bxv / wb - 0.5), ty = 2.0 * (byv / wb - 0.5);
            double px, py;
            H.project(tx, ty, px, py);
            // truncation toward zero; no sample where the projection is not a number
            const bool finite = px == px && py == py && fabs(px) < 2e9 && fabs(py) < 2e9;
            const int ix = finite ? (int)px : -1, iy = finite ? (int)py : -1;
            int kind = 0;
            double v = 0.0;
            if (ix >= 0 && iy >= 0 && ix < (int)w && iy < (int)h) {
                v = gray[(size_t)iy * w + ix];
                kind = white ? 1 : 2;
            }
            sx[s] = tx;
            sy[s] = ty;
            sv[s] = v;
            skind[s] = kind;
        }
        for (int k = lane; k < tw * tw; k += 64) vals[k] = 0.0f;
        __syncthreads();
        if (lane == 0) {
            double a[2][6] = {}, b[2][3] = {};
            for (int s = 0; s < ns; s++) {
                if (!skind[s]) continue;
                const int m = skind[s] - 1;
                const double x = sx[s], y = sy[s], v = sv[s];
                a[m][0] += x * x;
                a[m][1] += x * y;
                a[m][2] += x;
                a[m][3] += y * y;
                a[m][4] += y;
                a[m][5] += 1;
                b[m][0] += x * v;
                b[m][1] += y * v;
                b[m][2] += v;
            }
            double cw[3], cb[3];
            const bool solved = solve_plane(a[0], b[0], cw) && solve_plane(a[1], b[1], cb);
            for (int k = 0; k < 3; k++) {
                model[k] = cw[k];
                model[3 + k] = cb[k];
            }
            polarity_ok = (solved && (cw[2] - cb[2] < 0.0) == (L.rev != 0)) ? 1 : 0;
        }
        __syncthreads();
        if (!polarity_ok) {
            if (lane == 0) {
                atq_tag_code o{};
                o.code = 0;
                o.decision_margin = -1.0f;
                for (int k = 0; k < 9; k++) o.homography[k] = (float)H.m[k];
                out[q] = o;
            }
            __syncthreads();
            continue;
        }
        const double cw[3] = {model[0], model[1], model[2]}, cb[3] = {model[3], model[4], model[5]};
        for (int i = lane; i < L.nbits; i += 64) {
            const int bx = L.bx[i], by = L.by[i];
            const double tx = 2.0 * ((bx + 0.5) / wb - 0.5), ty = 2.0 * ((by + 0.5) / wb - 0.5);
            double px, py;
            H.project(tx, ty, px, py);
            const double v = bilinear(gray, w, h, px, py);
            if (v == -1.0) continue;
            const double th = (plane_at(cb, tx, ty) + plane_at(cw, tx, ty)) / 2.0;
            vals[(by - mn) * tw + bx - mn] = (float)(v - th);
        }
        __syncthreads();
        for (int k = lane; k < tw * tw; k += 64) {
            const int y = k / tw, x = k % tw;
            float acc = 0.0f;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const int yy = y + i - 1, xx = x + j - 1;
                    if (yy < 0 || yy >= tw || xx < 0 || xx >= tw) continue;
                    const float kv = (i == 1 && j == 1) ? 4.0f : ((i == 1) != (j == 1)) ? -1.0f : 0.0f;
                    acc = acc + vals[yy * tw + xx] * kv;
                }
            sharp[k] = acc;
        }
        __syncthreads();
        for (int k = lane; k < tw * tw; k += 64)
            sharp[k] = (float)((double)vals[k] + (double)L.sharp * (double)sharp[k]);
        __syncthreads();
        if (lane == 0) {
            uint64_t code = 0;
            float ws = 0.0f, bs = 0.0f, wc = 1.0f, bc = 1.0f;
            for (int i = 0; i < L.nbits; i++) {
                const float v = sharp[(L.by[i] - mn) * tw + L.bx[i] - mn];
                code <<= 1;
                if (v > 0.0f) {
                    ws += v;
                    wc += 1.0f;
                    code |= 1;
                } else {
                    bs -= v;
                    bc += 1.0f;
                }
            }
            atq_tag_code o{};
            o.code = code;
            o.decision_margin = fminf(ws / wc, bs / bc);
            for (int k = 0; k < 9; k++) o.homography[k] = (float)H.m[k];
            out[q] = o;
        }
        __syncthreads();
    }
}

struct Rotations {
    int32_t nbits;
    uint8_t perm[4][ATQ_TAG_MAX_BITS];  // bit i of the read code -> bit perm[r][i]
};

ATQ_DEV uint64_t rotate_code(uint64_t c, const Rotations &R, int r) {
    uint64_t o = 0;
    const int n = R.nbits;
    for (int i = 0; i < n; i++)
        if ((c >> (n - 1 - i)) & 1ull) o |= 1ull << (n - 1 - R.perm[r][i]);
    return o;
}

__global__ void __launch_bounds__(256)
k_match(const atq_tag_code *__restrict__ codes, uint32_t n_cap, const uint32_t *n_dev,
        const uint64_t *__restrict__ table, uint32_t n_table, Rotations R, int max_hamming,
        atq_tag_match *__restrict__ out) {
    const uint32_t n = n_dev ? min(*n_dev, n_cap) : n_cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        atq_tag_match m{-1, 0, 0, 0};
        const atq_tag_code c = codes[i];
        if (c.decision_margin >= 0.0f) {
            int best = max_hamming + 1, id = -1, rot = 0;
            for (int r = 0; r < 4; r++) {
                const uint64_t cr = rotate_code(c.code, R, r);
                for (uint32_t j = 0; j < n_table; j++) {
                    const int d = __popcll(cr ^ table[j]);
                    if (d < best) {
                        best = d;
                        id = (int)j;
                        rot = r;
                    }
                }
            }
            if (id >= 0) m = atq_tag_match{id, best, rot, 0};
        }
        out[i] = m;
    }
}

bool layout_ok(const atq_tag_layout &l) {
    if (l.width_at_border < 2 || l.width_at_border > ATQ_TAG_MAX_WIDTH) return false;
    if (l.total_width < l.width_at_border || l.total_width > ATQ_TAG_MAX_WIDTH) return false;
    if ((l.total_width - l.width_at_border) & 1) return false;
    if (l.nbits < 1 || l.nbits > ATQ_TAG_MAX_BITS) return false;
    const int mn = (l.width_at_border - l.total_width) / 2;
    for (int i = 0; i < l.nbits; i++) {
        const int x = l.bit_x[i] - mn, y = l.bit_y[i] - mn;
        if (x < 0 || y < 0 || x >= l.total_width || y >= l.total_width) return false;
    }
    return true;
}

}  // namespace

int launch_decode_payloads(hipStream_t s, const uint8_t *gray, size_t w, size_t h,
                           const atq_quad_corners *quads, size_t n, const uint32_t *n_dev,
                           const atq_tag_layout &l, atq_tag_code *out) {
    if (!layout_ok(l)) return ATQ_ERR_BAD_ARG;
    DecodeLayout L{};
    L.wb = l.width_at_border;
    L.tw = l.total_width;
    L.rev = l.reversed_border;
    L.nbits = l.nbits;
    L.sharp = l.decode_sharpening;
    for (int i = 0; i < ATQ_TAG_MAX_BITS; i++) {
        L.bx[i] = l.bit_x[i];
        L.by[i] = l.bit_y[i];
    }
    if (n == 0) return ATQ_OK;
    const uint32_t grid = (uint32_t)std::min<size_t>(n, 65536);
    k_decode<<<grid, 64, 0, s>>>(gray, (uint32_t)w, (uint32_t)h, quads, (uint32_t)n, n_dev, L, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_match_codes(hipStream_t s, const atq_tag_code *codes, size_t n, const uint32_t *n_dev,
                       const uint64_t *table, size_t n_table, const atq_tag_layout &l,
                       int max_hamming, atq_tag_match *out) {
    if (!layout_ok(l) || max_hamming < 0 || n_table > 0xFFFFFFFFu) return ATQ_ERR_BAD_ARG;
    // the quarter turn (x, y) -> (wb - 1 - y, x) applied r times, as a bit permutation
    Rotations R{};
    R.nbits = l.nbits;
    const int wb = l.width_at_border;
    for (int r = 0; r < 4; r++)
        for (int i = 0; i < l.nbits; i++) {
            int x = l.bit_x[i], y = l.bit_y[i];
            for (int k = 0; k < r; k++) {
                const int nx = wb - 1 - y, ny = x;
                x = nx;
                y = ny;
            }
            int at = -1;
            for (int j = 0; j < l.nbits; j++)
                if (l.bit_x[j] == x && l.bit_y[j] == y) at = j;
            if (at < 0) return ATQ_ERR_BAD_ARG;
            R.perm[r][i] = (uint8_t)at;
        }
    if (n == 0) return ATQ_OK;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    k_match<<<grid, 256, 0, s>>>(codes, (uint32_t)n, n_dev, table, (uint32_t)n_table, R,
                                 max_hamming, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
