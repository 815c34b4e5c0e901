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

// Grey value of an RGB(A) pixel: the reference reads its frames with
// stbi_load(..., STBI_grey) (src/main.cpp:265), whose conversion of 3- and
// 4-channel pixels is stbi__compute_y (include/stb_image.h:1834-1836; alpha
// is dropped).
ATQ_DEV uint32_t grey_of(const uint8_t *px) {
    return ((uint32_t)px[0] * 77u + (uint32_t)px[1] * 150u + (uint32_t)px[2] * 29u) >> 8;
}

// Decimated grey pixel (xx, yy) of frame F: the input pixel (f*xx, f*yy)
// (point subsample), converted when the input has colour channels.
ATQ_DEV uint32_t grey_at(const FrontFrame &F, uint32_t xx, uint32_t yy) {
    const uint8_t *p = F.in + ((size_t)(yy * F.f) * F.w_in + (size_t)xx * F.f) * F.channels;
    return F.channels == 1 ? (uint32_t)p[0] : grey_of(p);
}

// Point-subsample decimation (upstream apriltag image_u8_decimate for
// integer factors; absent from the reference, notes.txt:18-21), of grey or
// RGB(A) input.
__global__ void k_decimate(const uint8_t *__restrict__ in, uint32_t w_in, uint32_t f,
                           uint32_t w, uint32_t h, uint32_t channels, uint8_t *__restrict__ out) {
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint8_t *p = in + ((size_t)(y * f) * w_in + (size_t)x * f) * channels;
    out[(size_t)y * w + x] = (uint8_t)(channels == 1 ? (uint32_t)p[0] : grey_of(p));
}

// Decimated pixels x .. x+15 of row y (x a multiple of 16, y < h) as four
// little-endian words; pixels at or past w read 0.
ATQ_DEV void load_strip_row(const FrontFrame &F, bool vec_in, int x, int y, uint32_t (&wd)[4]) {
    const uint32_t w = F.w, w_in = F.w_in, f = F.f;
    if (F.channels != 1) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            uint32_t word = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int xx = x + 4 * q + j;
                if (xx < (int)w) word |= grey_at(F, (uint32_t)xx, (uint32_t)y) << (8 * j);
            }
            wd[q] = word;
        }
        return;
    }
    if (f == 1) {
        const uint8_t *row = F.in + (size_t)y * w_in;
        if (vec_in && x + 16 <= (int)w) {
            const uint4 v = *reinterpret_cast<const uint4 *>(row + x);
            wd[0] = v.x;
            wd[1] = v.y;
            wd[2] = v.z;
            wd[3] = v.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++)
                wd[q] = x + 4 * q < (int)w ? *reinterpret_cast<const uint32_t *>(row + x + 4 * q) : 0u;
        }
        return;
    }
    const uint8_t *row = F.in + (size_t)(y * f) * w_in;
    if (f == 2 && vec_in && 2 * x + 32 <= (int)w_in) {
        // even bytes of 32 input bytes
        const uint4 a = *reinterpret_cast<const uint4 *>(row + 2 * x);
        const uint4 b = *reinterpret_cast<const uint4 *>(row + 2 * x + 16);
        auto pack = [](uint32_t u, uint32_t v) {
            return (u & 0xFFu) | ((u >> 8) & 0xFF00u) | ((v & 0xFFu) << 16) | ((v & 0xFF0000u) << 8);
        };
        wd[0] = pack(a.x, a.y);
        wd[1] = pack(a.z, a.w);
        wd[2] = pack(b.x, b.y);
        wd[3] = pack(b.z, b.w);
        return;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t word = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int xx = x + 4 * q + j;
            if (xx < (int)w) word |= (uint32_t)row[(size_t)xx * f] << (8 * j);
        }
        wd[q] = word;
    }
}

// 0xFF in every byte of v greater than c (0 <= c <= 254, replicated in
// every byte of cc), 0 elsewhere, four bytes at a time: (v | H) - (c + 1)
// without H has the high bit of a byte set where its low 7 bits reach
// c + 1's; where the bytes' high bits differ v's decides instead.
ATQ_DEV uint32_t bytes_greater(uint32_t v, uint32_t cc) {
    constexpr uint32_t H = 0x80808080u;
    const uint32_t c1 = cc + 0x01010101u;
    const uint32_t t = (v | H) - (c1 & ~H);
    const uint32_t ge = ((t & ~(v ^ c1)) | (v & (v ^ c1))) & H;
    return (ge >> 7) * 0xFFu;
}

// Min and max of the bytes of four words (one 4x4 tile): the even and the
// odd bytes of each word as two 16-bit lanes, reduced with packed 16-bit
// min/max (v_pk_min_u16 / v_pk_max_u16) over the 8 lane pairs, then the
// two lanes.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
ATQ_DEV u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
ATQ_DEV void tile_minmax(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t &mn,
                         uint32_t &mx) {
    const uint32_t v4[4] = {a, b, c, d};
    u16x2 lo = as_u16x2(0x00FF00FFu), hi = as_u16x2(0u);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const u16x2 e = as_u16x2(v4[k] & 0x00FF00FFu), o = as_u16x2((v4[k] >> 8) & 0x00FF00FFu);
        lo = __builtin_elementwise_min(lo, __builtin_elementwise_min(e, o));
        hi = __builtin_elementwise_max(hi, __builtin_elementwise_max(e, o));
    }
    const uint32_t l = __builtin_bit_cast(uint32_t, lo), h = __builtin_bit_cast(uint32_t, hi);
    mn = min(l & 0xFFFFu, l >> 16);
    mx = max(h & 0xFFFFu, h >> 16);
}

// Streaming 16-byte loads and stores (the "nt" cache hint): the input
// pixels are read once and the thresholded bytes are written once here.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
ATQ_DEV uint4 load_stream(const uint8_t *p) {
    const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
ATQ_DEV void store_stream(uint8_t *p, uint4 v) {
    const u32x4 u = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(u, reinterpret_cast<u32x4 *>(p));
}

// Tile extrema of a region and its one-tile halo ring, min | max << 8
// (out-of-image tiles are neutral: min 255, max 0). Rows are padded to an
// even length so that a thread's 6 columns are 3 aligned words.
template <int RX, int RY>
struct alignas(16) ThresholdLds {
    uint16_t e[RY + 2][RX + 2];
};

// One region of RX x RY tiles of frame F. Each of the 256 threads owns a
// strip of 4 tiles (4 rows x 16 pixels) and keeps its pixels in registers:
// it loads them (threads < the ring size also load one halo tile), writes
// the strip's tile extrema to LDS, and after one barrier dilates its tiles
// over their 3x3 neighbourhoods and thresholds and stores its 64 pixels.
template <int RX, int RY>
struct RegionRegs {
    uint32_t px[4][4];  // the strip: 4 rows x 4 words
    uint32_t hp[4];     // the halo tile
    int tx0, ty0;       // region origin (tiles)
    bool live;
};

// FAST: grey input, no decimation, w a multiple of 16 and every buffer
// 16-byte aligned, no tile extrema asked for (the host checks this for
// every frame of the launch): each strip row is one streaming 16-byte load
// and store, with none of the general path's per-row branches.
template <int RX, int RY, bool FAST>
ATQ_DEV void region_load(RegionRegs<RX, RY> &R, const FrontFrame &F, uint32_t region,
                         uint32_t regions_x, bool vec_in, bool live) {
    constexpr int SX = RX / 4;
    static_assert(RX % 4 == 0 && SX * RY == 256, "one 4-tile strip per thread");
    constexpr int RING = 2 * (RX + 2) + 2 * RY;
    static_assert(RING <= 256, "one halo tile per thread");
    const int tw = (int)(F.w / 4), th = (int)(F.h / 4);
    const int tid = threadIdx.x;
    R.live = live;
    R.tx0 = (int)(region % regions_x) * RX;
    R.ty0 = (int)(region / regions_x) * RY;
    const int sx = tid % SX, sy = tid / SX;
    const int ty = R.ty0 + sy, tx = R.tx0 + 4 * sx;  // the strip's first tile
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) R.px[r][q] = 0;
    if (FAST) {
        // w % 16 == 0: a strip that starts in the frame ends in it
        if (live && ty < th && tx < tw) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint4 v = load_stream(F.in + (size_t)(4 * ty + r) * F.w + 4 * tx);
                R.px[r][0] = v.x;
                R.px[r][1] = v.y;
                R.px[r][2] = v.z;
                R.px[r][3] = v.w;
            }
        }
    } else if (live && ty < th) {
#pragma unroll
        for (int r = 0; r < 4; r++) load_strip_row(F, vec_in, 4 * tx, 4 * ty + r, R.px[r]);
    }
    int hx = 0, hy = 0;  // halo tile (region-relative, -1 .. RX / RY)
    if (tid < RING) {
        if (tid < RX + 2) { hx = tid - 1; hy = -1; }
        else if (tid < 2 * (RX + 2)) { hx = tid - (RX + 2) - 1; hy = RY; }
        else if (tid < 2 * (RX + 2) + RY) { hx = -1; hy = tid - 2 * (RX + 2); }
        else { hx = RX; hy = tid - 2 * (RX + 2) - RY; }
    }
    const int htx = R.tx0 + hx, hty = R.ty0 + hy;
    const bool halo_in = live && tid < RING && htx >= 0 && htx < tw && hty >= 0 && hty < th;
#pragma unroll
    for (int r = 0; r < 4; r++) R.hp[r] = 0;
    if (FAST && halo_in) {
#pragma unroll
        for (int r = 0; r < 4; r++)
            R.hp[r] = *reinterpret_cast<const uint32_t *>(F.in + (size_t)(4 * hty + r) * F.w + 4 * htx);
    } else if (halo_in) {
        const uint32_t f = F.f;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint8_t *row = F.in + (size_t)((4 * hty + r) * f) * F.w_in;
            if (F.channels != 1) {
#pragma unroll
                for (int j = 0; j < 4; j++)
                    R.hp[r] |= grey_at(F, (uint32_t)(4 * htx + j), (uint32_t)(4 * hty + r)) << (8 * j);
            } else if (f == 1) {
                R.hp[r] = *reinterpret_cast<const uint32_t *>(row + 4 * htx);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) R.hp[r] |= (uint32_t)row[(size_t)(4 * htx + j) * f] << (8 * j);
            }
        }
    }
}

// Tile extrema of the strip and of the halo tile to LDS (out-of-image
// tiles are neutral: min 255, max 0); the raw extrema to F.ext if asked.
template <int RX, int RY, bool FAST>
ATQ_DEV void region_extrema(ThresholdLds<RX, RY> &L, const RegionRegs<RX, RY> &R,
                            const FrontFrame &F) {
    constexpr int SX = RX / 4;
    constexpr int RING = 2 * (RX + 2) + 2 * RY;
    const int tw = (int)(F.w / 4), th = (int)(F.h / 4);
    const int tid = threadIdx.x;
    const int sx = tid % SX, sy = tid / SX;
    const int ty = R.ty0 + sy, tx = R.tx0 + 4 * sx;
    const bool row_in = R.live && ty < th;
    uint32_t smn[4], smx[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        smn[q] = 255;
        smx[q] = 0;
        if (row_in && tx + q < tw) tile_minmax(R.px[0][q], R.px[1][q], R.px[2][q], R.px[3][q], smn[q], smx[q]);
        L.e[sy + 1][4 * sx + q + 1] = (uint16_t)(smn[q] | (smx[q] << 8));
    }
    if (!FAST && F.ext && row_in) {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (tx + q < tw)
                F.ext[(size_t)ty * tw + tx + q] = atq_extrema{(uint8_t)smn[q], (uint8_t)smx[q]};
    }
    if (tid < RING) {
        int hx, hy;
        if (tid < RX + 2) { hx = tid - 1; hy = -1; }
        else if (tid < 2 * (RX + 2)) { hx = tid - (RX + 2) - 1; hy = RY; }
        else if (tid < 2 * (RX + 2) + RY) { hx = -1; hy = tid - 2 * (RX + 2); }
        else { hx = RX; hy = tid - 2 * (RX + 2) - RY; }
        const int htx = R.tx0 + hx, hty = R.ty0 + hy;
        uint32_t mn = 255, mx = 0;
        if (R.live && htx >= 0 && htx < tw && hty >= 0 && hty < th)
            tile_minmax(R.hp[0], R.hp[1], R.hp[2], R.hp[3], mn, mx);
        L.e[hy + 1][hx + 1] = (uint16_t)(mn | (mx << 8));
    }
}

// After the barrier: 3x3 dilation of the strip's tiles, then the threshold
// of its pixels (127 if max-min < min_diff, else 255 if
// v > min+(max-min)/2, else 0), stored with the decimated grey when f > 1.
template <int RX, int RY, bool FAST>
ATQ_DEV void region_finish(const ThresholdLds<RX, RY> &L, const RegionRegs<RX, RY> &R,
                           const FrontFrame &F, bool vec_out) {
    constexpr int SX = RX / 4;
    const uint32_t w = F.w;
    const int tw = (int)(F.w / 4), th = (int)(F.h / 4);
    const int tid = threadIdx.x;
    const int sx = tid % SX, sy = tid / SX;
    const int ty = R.ty0 + sy, tx = R.tx0 + 4 * sx;
    const bool row_in = R.live && ty < th;
    // the 3 x 6 neighbourhood of the strip's tiles (LDS columns 4 sx ..
    // 4 sx + 5) as 3 aligned words per row
    uint32_t nb[3][6];
#pragma unroll
    for (int dr = 0; dr < 3; dr++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(&L.e[sy + dr][4 * sx + 2 * k]);
            nb[dr][2 * k] = v & 0xFFFFu;
            nb[dr][2 * k + 1] = v >> 16;
        }
    uint32_t out[4][4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint32_t mn = 255, mx = 0;
#pragma unroll
        for (int dr = 0; dr < 3; dr++)
#pragma unroll
            for (int dc = 0; dc < 3; dc++) {
                mn = min(mn, nb[dr][q + dc] & 0xFFu);
                mx = max(mx, nb[dr][q + dc] >> 8);
            }
        if (!FAST && F.fil && row_in && tx + q < tw)
            F.fil[(size_t)ty * tw + tx + q] = atq_extrema{(uint8_t)mn, (uint8_t)mx};
        const bool flat = (int)(mx - mn) < F.min_diff;
        const uint32_t comp = (mn + (mx - mn) / 2) & 0xFFu;
        // no byte exceeds 255 (comp == 255 only when min == max == 255)
        const uint32_t cc = comp * 0x01010101u;
#pragma unroll
        for (int r = 0; r < 4; r++)
            out[r][q] = flat ? 0x7F7F7F7Fu : (comp < 255 ? bytes_greater(R.px[r][q], cc) : 0u);
    }
    if (FAST) {
        if (row_in && tx < tw) {
#pragma unroll
            for (int r = 0; r < 4; r++)
                store_stream(F.thr + (size_t)(4 * ty + r) * w + 4 * tx,
                             make_uint4(out[r][0], out[r][1], out[r][2], out[r][3]));
        }
    } else if (row_in && tx < tw) {
        uint8_t *gray_out = F.f > 1 || F.channels != 1 ? F.gray : nullptr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const size_t o = (size_t)(4 * ty + r) * w + 4 * tx;
            if (vec_out && tx + 4 <= tw) {
                *reinterpret_cast<uint4 *>(F.thr + o) = make_uint4(out[r][0], out[r][1], out[r][2], out[r][3]);
                if (gray_out)
                    *reinterpret_cast<uint4 *>(gray_out + o) =
                        make_uint4(R.px[r][0], R.px[r][1], R.px[r][2], R.px[r][3]);
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    if (tx + q >= tw) break;
                    *reinterpret_cast<uint32_t *>(F.thr + o + 4 * q) = out[r][q];
                    if (gray_out) *reinterpret_cast<uint32_t *>(gray_out + o + 4 * q) = R.px[r][q];
                }
            }
        }
    }
}

// Persistent: the work-groups walk the regions of every frame of the table
// in turn (XCD-contiguous runs of regions within a frame). A frame's
// regions are normally all resident at once, so that all of its loads are
// in flight together, which HBM's latency-bandwidth product asks for.
template <int RX, int RY, bool FAST>
__global__ void __launch_bounds__(256)
k_threshold(FrontTable tab) {
    __shared__ ThresholdLds<RX, RY> L;
    for (uint32_t fi = 0; fi < tab.n; fi++) {
        const FrontFrame &F = tab.fr[fi];
        const uint32_t regions_x = (F.w / 4 + RX - 1) / RX;
        const uint32_t n_regions = regions_x * ((F.h / 4 + RY - 1) / RY);
        const bool vec_in = F.channels == 1 && (F.w_in & 15) == 0 &&
                            (reinterpret_cast<uintptr_t>(F.in) & 15) == 0;
        const bool vec_out = (F.w & 15) == 0 && (reinterpret_cast<uintptr_t>(F.thr) & 15) == 0 &&
                             (reinterpret_cast<uintptr_t>(F.gray) & 15) == 0;
        for (uint32_t it = blockIdx.x; it < xcd_grid(n_regions); it += gridDim.x) {
            const uint32_t region = xcd_tile(it, n_regions);
            if (region >= n_regions) continue;
            RegionRegs<RX, RY> a;
            region_load<RX, RY, FAST>(a, F, region, regions_x, vec_in, true);
            region_extrema<RX, RY, FAST>(L, a, F);
            __syncthreads();
            region_finish<RX, RY, FAST>(L, a, F, vec_out);
            __syncthreads();  // the LDS is reused by the next region
        }
    }
}

int launch_decimate(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                    uint8_t *out, int channels) {
    size_t w = 1 + (w_in - 1) / f, h = 1 + (h_in - 1) / f;
    dim3 grid(atq_div_up(w, 256), (unsigned)h);
    k_decimate<<<grid, 256, 0, s>>>(in, (uint32_t)w_in, (uint32_t)f, (uint32_t)w, (uint32_t)h,
                                    (uint32_t)channels, out);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

bool front_shapes_supported(const FrontShapes &sh) {
    const bool thr = (sh.thr_rx == 64 && sh.thr_ry == 16) || (sh.thr_rx == 32 && sh.thr_ry == 32) ||
                     (sh.thr_rx == 16 && sh.thr_ry == 64);
    return thr && (sh.ccl_rows == 8 || sh.ccl_rows == 16 || sh.ccl_rows == 32);
}

int front_shape_candidates(FrontShapes *out, int max_out) {
    static const int thr[3][2] = {{64, 16}, {32, 32}, {16, 64}};
    int n = 0;
    for (int c : {16, 8, 32})
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
        const uint32_t n = rx * ry;
        items = std::max(items, xcd_grid(n));
    }
    return items;
}

// Every frame of the launch can take the FAST path (region_load).
bool threshold_fast(const FrontTable &tab) {
    for (uint32_t i = 0; i < tab.n; i++) {
        const FrontFrame &F = tab.fr[i];
        const bool aligned = ((reinterpret_cast<uintptr_t>(F.in) | reinterpret_cast<uintptr_t>(F.thr)) & 15) == 0;
        if (F.channels != 1 || F.f != 1 || F.w != F.w_in || (F.w & 15) || !aligned || F.ext || F.fil)
            return false;
    }
    return true;
}

template <int RX, int RY>
void threshold_launch(hipStream_t s, const FrontTable &tab) {
    // persistent: at most 2048 work-groups walk every frame's regions
    const uint32_t grid = std::min<uint32_t>(threshold_items<RX, RY>(tab), 2048u);
    if (threshold_fast(tab))
        k_threshold<RX, RY, true><<<grid ? grid : 8, 256, 0, s>>>(tab);
    else
        k_threshold<RX, RY, false><<<grid ? grid : 8, 256, 0, s>>>(tab);
}
}  // namespace

int launch_threshold_frames(hipStream_t s, const FrontFrame *frames, int n, const FrontShapes &sh) {
    if (n < 1 || n > FRONT_MAX_FRAMES || !front_shapes_supported(sh)) return ATQ_ERR_BAD_ARG;
    FrontTable tab{};
    for (int i = 0; i < n; i++) tab.fr[i] = frames[i];
    tab.n = (uint32_t)n;
    if (sh.thr_rx == 64) threshold_launch<64, 16>(s, tab);
    else if (sh.thr_rx == 32) threshold_launch<32, 32>(s, tab);
    else threshold_launch<16, 64>(s, tab);
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
    fr.channels = 1;
    return launch_threshold_frames(s, &fr, 1, FrontShapes{});
}

}  // namespace atq
