// kernels.hpp - launchers of the gfx950 kernels (one .hip file per stage).
#pragma once
#include "common.hpp"

namespace atq {

struct ScanCtx;  // scan.hpp

// threshold.hip
int launch_decimate(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                    uint8_t *out);
int launch_threshold(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                     uint8_t *gray_out, atq_extrema *ext, atq_extrema *fil, uint8_t *thr,
                     int min_diff);

// ccl.hip
int launch_ccl(hipStream_t s, const uint8_t *thr, size_t w, size_t h, uint32_t *parent,
               uint32_t *labels, uint32_t *sizes, uint32_t size_cap = 0xFFFFFFFFu);

// boundary.hip
int launch_find_boundaries_sparse(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                  size_t w, size_t h, atq_boundary_point *points,
                                  uint64_t *keys, int min_size);
// The detector's open-addressing table of cluster pairs and record counts
// (slots = mask + 1, a power of two; zeroed before each frame).
struct PairTable {
    uint64_t *keys;
    uint32_t *cnt;
    uint32_t mask;
};
// Records of the pairs whose count lies in [24, 2 * (2w + 2h)], packed keys
// ((min << lb | max) << 2 | plane), in the reference's order; tile_keys /
// tile_points hold 4 * w * h entries, tile_scratch 2 * tiles + 2.
int launch_find_boundaries_filtered(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                    size_t w, size_t h, int min_size, uint32_t label_bits,
                                    const PairTable &table, atq_counts *counts,
                                    uint64_t *tile_keys, atq_boundary_point *tile_points,
                                    uint32_t *tile_scratch, uint64_t *keys,
                                    atq_boundary_point *points, ScanCtx &sc);
constexpr unsigned BOUNDARY_TILE_PX = 1024;  // pixels per compaction tile

// diff.hip
int launch_diff_forward(hipStream_t s, const float *img, size_t w, size_t h,
                        const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                        size_t nq, float *corners, float *w_out);
int launch_diff_backward(hipStream_t s, const float *img, size_t w, size_t h,
                         const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                         size_t nq, const float *gcorners, float *gw, float *gimg);

}  // namespace atq
