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
// region shares with its neighbours mostly come from the same L2. The region
// shape (RX, RY) is chosen per resolution by the autotuner; one launch covers
// a table of frames, walked by a persistent grid.
#include <algorithm>

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
struct ThresholdLds {
    static constexpr int PW = (RX + 2) * 4;  // region width in pixels incl. halo
    static constexpr int PH = (RY + 2) * 4;
    static constexpr int TW = RX + 2, TH = RY + 2;
    uint32_t pix[PH][PW / 4];
    uint8_t emin[TH][TW], emax[TH][TW];
    uint8_t fmin[RY][RX], fmax[RY][RX];
};

// One region of frame F (all 256 threads): stage, extrema, dilation, output.
template <int RX, int RY>
ATQ_DEV void threshold_region(ThresholdLds<RX, RY> &L, const FrontFrame &F, uint32_t region,
                              uint32_t regions_x) {
    using Lds = ThresholdLds<RX, RY>;
    constexpr int PW = Lds::PW, PH = Lds::PH, TW = Lds::TW, TH = Lds::TH;
    const uint8_t *__restrict__ in = F.in;
    const uint32_t w_in = F.w_in, f = F.f, w = F.w, h = F.h;
    uint8_t *__restrict__ gray_out = f > 1 ? F.gray : nullptr;
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
        if (i < WORDS) L.pix[i / (PW / 4)][i % (PW / 4)] = words[k];
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
                uint32_t wv = L.pix[r * 4 + k][c];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    uint32_t v = (wv >> (8 * b)) & 0xFFu;
                    mn = min(mn, v);
                    mx = max(mx, v);
                }
            }
            if (F.ext && r >= 1 && r <= RY && c >= 1 && c <= RX)
                F.ext[(size_t)ty * tw + tx] = atq_extrema{(uint8_t)mn, (uint8_t)mx};
        }
        L.emin[r][c] = (uint8_t)mn;
        L.emax[r][c] = (uint8_t)mx;
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
                mn = min(mn, (uint32_t)L.emin[r + dr][c + dc]);
                mx = max(mx, (uint32_t)L.emax[r + dr][c + dc]);
            }
        L.fmin[r][c] = (uint8_t)mn;
        L.fmax[r][c] = (uint8_t)mx;
        int ty = ty0 + 1 + r, tx = tx0 + 1 + c;
        if (F.fil && ty < th && tx < tw)
            F.fil[(size_t)ty * tw + tx] = atq_extrema{(uint8_t)mn, (uint8_t)mx};
    }
    __syncthreads();

    // 4. threshold the interior: one word (4 pixels of one tile row) per item
    for (int i = tid; i < RY * 4 * RX; i += 256) {
        int r = i / RX, c = i % RX;          // pixel row in interior, tile column
        int y = py0 + 4 + r, x = px0 + 4 + c * 4;
        if (y >= (int)h || x >= (int)w) continue;
        uint32_t wv = L.pix[4 + r][1 + c];
        uint32_t mn = L.fmin[r / 4][c], mx = L.fmax[r / 4][c];
        uint32_t out;
        if ((int)(mx - mn) < F.min_diff) {
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
        *reinterpret_cast<uint32_t *>(F.thr + (size_t)y * w + x) = out;
        if (gray_out) *reinterpret_cast<uint32_t *>(gray_out + (size_t)y * w + x) = wv;
    }
    __syncthreads();  // the LDS is reused by the next region
}

// Persistent: the work-groups walk the regions of every frame of the table
// in turn (XCD-contiguous runs of regions within a frame).
template <int RX, int RY>
__global__ void __launch_bounds__(256)
k_threshold(FrontTable tab) {
    __shared__ ThresholdLds<RX, RY> L;
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        const uint32_t regions_x = (F.w / 4 + RX - 1) / RX;
        const uint32_t n_regions = regions_x * ((F.h / 4 + RY - 1) / RY);
        for (uint32_t it = blockIdx.x; it < xcd_grid(n_regions); it += gridDim.x) {
            const uint32_t region = xcd_tile(it, n_regions);
            if (region < n_regions) threshold_region<RX, RY>(L, F, region, regions_x);
        }
    }
}

int launch_decimate(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                    uint8_t *out) {
    size_t w = 1 + (w_in - 1) / f, h = 1 + (h_in - 1) / f;
    dim3 grid(atq_div_up(w, 256), (unsigned)h);
    k_decimate<<<grid, 256, 0, s>>>(in, (uint32_t)w_in, (uint32_t)f, (uint32_t)w, (uint32_t)h, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

bool front_shapes_supported(const FrontShapes &sh) {
    const bool thr = (sh.thr_rx == 32 && sh.thr_ry == 16) || (sh.thr_rx == 64 && sh.thr_ry == 8) ||
                     (sh.thr_rx == 32 && sh.thr_ry == 8) || (sh.thr_rx == 16 && sh.thr_ry == 16);
    return thr && (sh.ccl_rows == 8 || sh.ccl_rows == 16);
}

int front_shape_candidates(FrontShapes *out, int max_out) {
    static const int thr[4][2] = {{32, 16}, {64, 8}, {32, 8}, {16, 16}};
    int n = 0;
    for (int c : {16, 8})
        for (const auto &t : thr) {
            if (n < max_out) out[n] = FrontShapes{t[0], t[1], c};
            n++;
        }
    return n;
}

namespace {
template <int RX, int RY>
uint32_t threshold_items(const FrontTable &tab) {
    uint32_t items = 0;
    for (uint32_t i = 0; i < tab.n; i++) {
        const uint32_t rx = (tab.fr[i].w / 4 + RX - 1) / RX, ry = (tab.fr[i].h / 4 + RY - 1) / RY;
        items = std::max(items, xcd_grid(rx * ry));
    }
    return items;
}

template <int RX, int RY>
void threshold_launch(hipStream_t s, const FrontTable &tab) {
    // persistent: at most 2048 work-groups walk every frame's regions
    const uint32_t grid = std::min<uint32_t>(threshold_items<RX, RY>(tab), 2048u);
    k_threshold<RX, RY><<<grid ? grid : 8, 256, 0, s>>>(tab);
}
}  // namespace

int launch_threshold_frames(hipStream_t s, const FrontFrame *frames, int n, const FrontShapes &sh) {
    if (n < 1 || n > FRONT_MAX_FRAMES || !front_shapes_supported(sh)) return ATQ_ERR_BAD_ARG;
    FrontTable tab{};
    for (int i = 0; i < n; i++) tab.fr[i] = frames[i];
    tab.n = (uint32_t)n;
    if (sh.thr_rx == 32 && sh.thr_ry == 16) threshold_launch<32, 16>(s, tab);
    else if (sh.thr_rx == 64) threshold_launch<64, 8>(s, tab);
    else if (sh.thr_rx == 32) threshold_launch<32, 8>(s, tab);
    else threshold_launch<16, 16>(s, tab);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_threshold(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                     uint8_t *gray_out, atq_extrema *ext, atq_extrema *fil, uint8_t *thr,
                     int min_diff) {
    FrontFrame fr{};
    fr.in = in;
    fr.gray = gray_out;
    fr.ext = ext;
    fr.fil = fil;
    fr.thr = thr;
    fr.w_in = (uint32_t)w_in;
    fr.h_in = (uint32_t)h_in;
    fr.f = (uint32_t)(f > 1 ? f : 1);
    fr.w = (uint32_t)(f > 1 ? 1 + (w_in - 1) / f : w_in);
    fr.h = (uint32_t)(f > 1 ? 1 + (h_in - 1) / f : h_in);
    fr.min_diff = min_diff;
    return launch_threshold_frames(s, &fr, 1, FrontShapes{});
}

}  // namespace atq
