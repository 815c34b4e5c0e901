// threshold.hip - decimation + adaptive tile threshold (src/threshold.cpp).
//
// One kernel does what the reference does in three passes
// (src/threshold.cpp:8-95): a work-group stages a region of RX x RY 4x4
// tiles plus a one-tile halo in LDS (reading the input once, decimating on
// the fly), computes every tile's {min,max}, dilates them over the 3x3 tile
// neighbourhood (edge-clipped), and thresholds the interior pixels:
//   127 if max-min < min_diff, else 255 if v > min+(max-min)/2, else 0.
// HBM traffic: f*Npx read (rows touched) + Npx written (+ Npx decimated grey
// when f > 1); regions go to the XCDs in contiguous runs, so the halo lines a
// region shares with its neighbours mostly come from the same L2.
#include "kernels.hpp"

namespace atq {

// Point-subsample decimation (upstream apriltag image_u8_decimate for
// integer factors; absent from the reference, notes.txt:18-21).
__global__ void k_decimate(const uint8_t *__restrict__ in, uint32_t w_in, uint32_t f,
                           uint32_t w, uint32_t h, uint8_t *__restrict__ out) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y;
    if (x >= w || y >= h) return;
    out[(size_t)y * w + x] = in[(size_t)(y * f) * w_in + (size_t)x * f];
}

template <int RX, int RY>
__global__ void __launch_bounds__(256)
k_threshold(const uint8_t *__restrict__ in, uint32_t w_in, uint32_t f, uint32_t w, uint32_t h,
            uint32_t regions_x, uint32_t n_regions, uint8_t *__restrict__ gray_out,
            atq_extrema *__restrict__ ext_out, atq_extrema *__restrict__ fil_out,
            uint8_t *__restrict__ thr_out, int min_diff) {
    constexpr int PW = (RX + 2) * 4;  // region width in pixels incl. halo
    constexpr int PH = (RY + 2) * 4;
    constexpr int TW = RX + 2, TH = RY + 2;
    __shared__ uint32_t pix[PH][PW / 4];
    __shared__ uint8_t emin[TH][TW], emax[TH][TW];
    __shared__ uint8_t fmin[RY][RX], fmax[RY][RX];

    const uint32_t region = xcd_tile(blockIdx.x, n_regions);
    if (region >= n_regions) return;
    const int tw = (int)(w / 4), th = (int)(h / 4);
    const int tx0 = (int)(region % regions_x) * RX - 1;  // first (halo) tile of the region
    const int ty0 = (int)(region / regions_x) * RY - 1;
    const int px0 = tx0 * 4, py0 = ty0 * 4;
    const int tid = threadIdx.x;

    // 1. stage decimated pixels of the region (words of 4 pixels); every
    // load of a thread is issued before the first LDS store
    constexpr int WORDS = PH * (PW / 4);
    constexpr int PER = (WORDS + 255) / 256;
    uint32_t words[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + 256 * k;
        const int r = i / (PW / 4), c = i % (PW / 4);
        const int y = py0 + r, x = px0 + c * 4;
        uint32_t word = 0;
        if (i < WORDS && y >= 0 && y < (int)h && x >= 0 && x < (int)w) {
            if (f == 1) {
                word = *reinterpret_cast<const uint32_t *>(in + (size_t)y * w_in + x);
            } else if (f == 2 && (w_in & 3) == 0) {
                const uint32_t *row = reinterpret_cast<const uint32_t *>(in + (size_t)(2 * y) * w_in);
                uint32_t a = row[(2 * x) / 4], b = row[(2 * x) / 4 + 1];
                word = (a & 0xFFu) | ((a >> 8) & 0xFF00u) | ((b & 0xFFu) << 16) | ((b & 0xFF0000u) << 8);
            } else {
                const uint8_t *row = in + (size_t)(y * f) * w_in;
#pragma unroll
                for (int q = 0; q < 4; q++) word |= (uint32_t)row[(size_t)(x + q) * f] << (8 * q);
            }
        }
        words[k] = word;
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = tid + 256 * k;
        if (i < WORDS) pix[i / (PW / 4)][i % (PW / 4)] = words[k];
    }
    __syncthreads();

    // 2. raw tile extrema (out-of-image tiles are neutral: min 255, max 0)
    for (int i = tid; i < TH * TW; i += 256) {
        int r = i / TW, c = i % TW;
        int ty = ty0 + r, tx = tx0 + c;
        uint32_t mn = 255, mx = 0;
        if (ty >= 0 && ty < th && tx >= 0 && tx < tw) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint32_t wv = pix[r * 4 + k][c];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    uint32_t v = (wv >> (8 * b)) & 0xFFu;
                    mn = min(mn, v);
                    mx = max(mx, v);
                }
            }
            if (ext_out && r >= 1 && r <= RY && c >= 1 && c <= RX)
                ext_out[(size_t)ty * tw + tx] = atq_extrema{(uint8_t)mn, (uint8_t)mx};
        }
        emin[r][c] = (uint8_t)mn;
        emax[r][c] = (uint8_t)mx;
    }
    __syncthreads();

    // 3. 3x3 dilation of the interior tiles
    for (int i = tid; i < RY * RX; i += 256) {
        int r = i / RX, c = i % RX;
        uint32_t mn = 255, mx = 0;
#pragma unroll
        for (int dr = 0; dr < 3; dr++)
#pragma unroll
            for (int dc = 0; dc < 3; dc++) {
                mn = min(mn, (uint32_t)emin[r + dr][c + dc]);
                mx = max(mx, (uint32_t)emax[r + dr][c + dc]);
            }
        fmin[r][c] = (uint8_t)mn;
        fmax[r][c] = (uint8_t)mx;
        int ty = ty0 + 1 + r, tx = tx0 + 1 + c;
        if (fil_out && ty < th && tx < tw)
            fil_out[(size_t)ty * tw + tx] = atq_extrema{(uint8_t)mn, (uint8_t)mx};
    }
    __syncthreads();

    // 4. threshold the interior: one word (4 pixels of one tile row) per item
    for (int i = tid; i < RY * 4 * RX; i += 256) {
        int r = i / RX, c = i % RX;          // pixel row in interior, tile column
        int y = py0 + 4 + r, x = px0 + 4 + c * 4;
        if (y >= (int)h || x >= (int)w) continue;
        uint32_t wv = pix[4 + r][1 + c];
        uint32_t mn = fmin[r / 4][c], mx = fmax[r / 4][c];
        uint32_t out;
        if ((int)(mx - mn) < min_diff) {
            out = 0x7F7F7F7Fu;
        } else {
            uint32_t comp = (mn + (mx - mn) / 2) & 0xFFu;
            out = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                uint32_t v = (wv >> (8 * b)) & 0xFFu;
                out |= (v > comp ? 0xFFu : 0u) << (8 * b);
            }
        }
        *reinterpret_cast<uint32_t *>(thr_out + (size_t)y * w + x) = out;
        if (gray_out) *reinterpret_cast<uint32_t *>(gray_out + (size_t)y * w + x) = wv;
    }
}

int launch_decimate(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                    uint8_t *out) {
    size_t w = 1 + (w_in - 1) / f, h = 1 + (h_in - 1) / f;
    dim3 grid(atq_div_up(w, 256), (unsigned)h);
    k_decimate<<<grid, 256, 0, s>>>(in, (uint32_t)w_in, (uint32_t)f, (uint32_t)w, (uint32_t)h, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_threshold(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                     uint8_t *gray_out, atq_extrema *ext, atq_extrema *fil, uint8_t *thr,
                     int min_diff) {
    constexpr int RX = 32, RY = 16;
    size_t w = f > 1 ? 1 + (w_in - 1) / f : w_in;
    size_t h = f > 1 ? 1 + (h_in - 1) / f : h_in;
    const uint32_t rx = atq_div_up(w / 4, RX), ry = atq_div_up(h / 4, RY);
    k_threshold<RX, RY><<<xcd_grid(rx * ry), 256, 0, s>>>(
        in, (uint32_t)w_in, (uint32_t)f, (uint32_t)w, (uint32_t)h, rx, rx * ry,
        f > 1 ? gray_out : nullptr, ext, fil, thr, min_diff);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
