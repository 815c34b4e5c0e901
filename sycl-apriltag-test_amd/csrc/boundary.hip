// boundary.hip - boundary-point extraction (row A7, src/find_boundaries.cpp).
//
// For a pixel (x >= 1; column 0 is never a local pixel in the reference's
// 8x32 work-group tiling, src/find_boundaries.cpp:50-92) with label L, the
// neighbours RIGHT (0,+1), BOTTOM_LEFT (+1,-1), BOTTOM (+1,0) and
// BOTTOM_RIGHT (+1,+1) are tested (out-of-image neighbours read label 0).
// A record is emitted when the white bits differ and both components have
// size >= min_size; BOTTOM_LEFT is skipped when the left and down neighbours
// differ in white bit and both are large (src/find_boundaries.cpp:148-162).
// Record: packed_x = x | plane<<14, packed_y = y | (neighbour white and
// local not) << 15; key = min(label values) << 32 | max(label values).
//
// Two forms: the drop-in writes the reference's 4 sparse planes in full;
// the pipeline form writes the records densely in the reference's order.
#include "kernels.hpp"
#include "scan.hpp"

namespace atq {

namespace {

struct PixelRecords {
    uint32_t n;
    atq_boundary_point p[4];
    uint64_t k[4];
    uint8_t plane[4];
};

ATQ_DEV uint32_t label_at(const uint32_t *labels, uint32_t w, uint32_t h, uint32_t x, uint32_t y) {
    return (x < w && y < h) ? labels[(size_t)y * w + x] : 0u;
}

// key_shift: 32 for the reference's min<<32|max, lb for the packed form.
ATQ_DEV void pixel_records(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes,
                           uint32_t w, uint32_t h, uint32_t x, uint32_t y, uint32_t min_size,
                           uint32_t key_shift, PixelRecords &out) {
    out.n = 0;
    if (x == 0 || x >= w || y >= h) return;
    const uint32_t L = labels[(size_t)y * w + x];
    const uint32_t lv = L & LABEL_VALUE_MASK, wl = L & LABEL_PIXEL_MASK;
    const bool local_big = sizes[lv] >= min_size;
    const uint32_t right = label_at(labels, w, h, x + 1, y);
    const uint32_t down = label_at(labels, w, h, x, y + 1);
    const uint32_t down_left = label_at(labels, w, h, x - 1, y + 1);
    const uint32_t down_right = label_at(labels, w, h, x + 1, y + 1);
    auto test = [&](uint32_t T, uint32_t plane) {
        const uint32_t wt = T & LABEL_PIXEL_MASK;
        if (wt == wl) return;
        const uint32_t tv = T & LABEL_VALUE_MASK;
        if (!local_big || sizes[tv] < min_size) return;
        uint32_t i = out.n++;
        out.p[i].packed_x = (uint16_t)(x | (plane << 14));
        out.p[i].packed_y = (uint16_t)(wt > wl ? (COLOR_DIRECTION_MASK | y) : y);
        out.k[i] = ((uint64_t)min(lv, tv) << key_shift) | (uint64_t)max(lv, tv);
        out.plane[i] = (uint8_t)plane;
    };
    test(right, 0);
    {
        const uint32_t left = labels[(size_t)y * w + x - 1];
        bool dup = (left & LABEL_PIXEL_MASK) != (down & LABEL_PIXEL_MASK) &&
                   sizes[left & LABEL_VALUE_MASK] >= min_size &&
                   sizes[down & LABEL_VALUE_MASK] >= min_size;
        if (!dup) test(down_left, 1);
    }
    test(down, 2);
    test(down_right, 3);
}

}  // namespace

__global__ void __launch_bounds__(256)
k_boundaries_sparse(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes,
                    uint32_t w, uint32_t h, atq_boundary_point *__restrict__ points,
                    uint64_t *__restrict__ keys, uint32_t min_size) {
    uint32_t x = blockIdx.x * 64 + (threadIdx.x & 63);
    uint32_t y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    PixelRecords r;
    pixel_records(labels, sizes, w, h, x, y, min_size, 32, r);
    const size_t n = (size_t)w * h, lin = (size_t)y * w + x;
    atq_boundary_point pp[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    uint64_t kk[4] = {~0ull, ~0ull, ~0ull, ~0ull};
    for (uint32_t i = 0; i < r.n; i++) {
        pp[r.plane[i]] = r.p[i];
        kk[r.plane[i]] = r.k[i];
    }
#pragma unroll
    for (int p = 0; p < 4; p++) {
        points[p * n + lin] = pp[p];
        keys[p * n + lin] = kk[p];
    }
}

// Compact form, in the reference's stable order without atomics: tiles of
// 1024 consecutive pixels (4 per thread) count their records, the counts are
// scanned, and each tile writes its records at its offset in raster order.
// The key is the packed pair key (min << lb | max) with the plane in its two
// low bits, so a stable sort by it yields (pair, plane, y, x): the
// reference's plane-major compaction followed by its sort by pair key.
constexpr int BND_PX_PER_THREAD = 4;
constexpr int BND_TILE = 256 * BND_PX_PER_THREAD;

__global__ void __launch_bounds__(256)
k_bnd_count(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes, uint32_t w,
            uint32_t h, uint32_t min_size, uint32_t *__restrict__ tile_counts) {
    __shared__ uint32_t lds4[4];
    const size_t n = (size_t)w * h;
    const size_t p0 = (size_t)blockIdx.x * BND_TILE + threadIdx.x * BND_PX_PER_THREAD;
    uint32_t c = 0;
    if (p0 < n) {
        const uint32_t y = (uint32_t)(p0 / w), x0 = (uint32_t)(p0 % w);
        for (int k = 0; k < BND_PX_PER_THREAD; k++) {
            PixelRecords r;
            pixel_records(labels, sizes, w, h, x0 + k, y, min_size, 0, r);
            c += r.n;
        }
    }
    uint32_t t = block_reduce_sum<uint32_t>(c, lds4);
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = t;
}

__global__ void __launch_bounds__(256)
k_bnd_emit(const uint32_t *__restrict__ labels, const uint32_t *__restrict__ sizes, uint32_t w,
           uint32_t h, uint32_t min_size, uint32_t lb, const uint32_t *__restrict__ tile_off,
           atq_boundary_point *__restrict__ points, uint64_t *__restrict__ keys, uint32_t capacity,
           atq_counts *counts) {
    __shared__ uint32_t lds4[4];
    const size_t n = (size_t)w * h;
    const size_t p0 = (size_t)blockIdx.x * BND_TILE + threadIdx.x * BND_PX_PER_THREAD;
    const bool live = p0 < n;
    const uint32_t y = live ? (uint32_t)(p0 / w) : 0, x0 = live ? (uint32_t)(p0 % w) : 0;
    uint32_t c = 0;
    if (live)
        for (int k = 0; k < BND_PX_PER_THREAD; k++) {
            PixelRecords r;
            pixel_records(labels, sizes, w, h, x0 + k, y, min_size, 0, r);
            c += r.n;
        }
    uint32_t tot;
    uint32_t pos = tile_off[blockIdx.x] + block_exclusive_scan<uint32_t>(c, lds4, &tot);
    if (!live || c == 0) return;
    if (pos + c > capacity) {
        atomicOr(&counts->overflow, OVF_BOUNDARY);
        return;
    }
    for (int k = 0; k < BND_PX_PER_THREAD; k++) {
        PixelRecords r;
        pixel_records(labels, sizes, w, h, x0 + k, y, min_size, lb, r);
        for (uint32_t i = 0; i < r.n; i++, pos++) {
            points[pos] = r.p[i];
            keys[pos] = (r.k[i] << 2) | r.plane[i];
        }
    }
}

namespace {
struct TileCount {
    const uint32_t *c;
    ATQ_DEV uint32_t operator()(uint32_t i) const { return c[i]; }
};
struct TileOffset {
    uint32_t *off;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t) const { off[i] = pre; }
};
}  // namespace

int launch_find_boundaries_sparse(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                  size_t w, size_t h, atq_boundary_point *points, uint64_t *keys,
                                  int min_size) {
    dim3 grid(atq_div_up(w, 64), atq_div_up(h, 4));
    k_boundaries_sparse<<<grid, 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h, points, keys,
                                             (uint32_t)min_size);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

int launch_find_boundaries_compact(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                   size_t w, size_t h, atq_boundary_point *points, uint64_t *keys,
                                   uint32_t capacity, atq_counts *counts, int min_size,
                                   uint32_t label_bits, uint32_t *tile_scratch,
                                   uint32_t *scan_scratch) {
    const unsigned tiles = atq_div_up(w * h, BND_TILE);
    uint32_t *tile_counts = tile_scratch, *tile_off = tile_scratch + tiles;
    k_bnd_count<<<tiles, 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h, (uint32_t)min_size,
                                      tile_counts);
    int r = device_scan(s, nullptr, tiles, TileCount{tile_counts}, TileOffset{tile_off},
                        scan_scratch, &counts->n_boundary);
    if (r) return r;
    k_bnd_emit<<<tiles, 256, 0, s>>>(labels, sizes, (uint32_t)w, (uint32_t)h, (uint32_t)min_size,
                                     label_bits, tile_off, points, keys, capacity, counts);
    return hipGetLastError() == hipSuccess ? ATQ_OK : ATQ_ERR_HIP;
}

}  // namespace atq
