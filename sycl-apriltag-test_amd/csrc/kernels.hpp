// kernels.hpp - launchers of the gfx950 kernels (one .hip file per stage).
#pragma once
#include "common.hpp"

namespace atq {

struct ScanCtx;  // scan.hpp

// One frame of the front-end stages (decimate/threshold, CCL). Several
// frames go to one launch as a table passed by value; every kernel walks the
// frames in turn with a persistent grid.
struct FrontFrame {
    const uint8_t *in;     // input, w_in x h_in pixels of `channels` bytes
    uint8_t *gray;         // decimated grey (written when f > 1 or channels > 1; may be null)
    atq_extrema *ext, *fil;  // raw and dilated tile extrema (may be null)
    uint8_t *thr;          // thresholded, w x h
    uint32_t *parent, *labels, *sizes;
    uint32_t w_in, h_in, f, w, h;
    uint32_t size_cap;     // sizes saturate here (CCL)
    int32_t min_diff;
    uint32_t channels;     // 1 grey; 3 RGB or 4 RGBA, converted to grey on load
    // words the last CCL kernel zeroes for the later stages (16-byte
    // aligned; may be null), except clear[clear_idx] = clear_val
    uint32_t *clear;
    uint32_t clear_words, clear_idx, clear_val;
};

// The frame's clear (FrontFrame::clear), grid-stride, 16 bytes per store.
ATQ_DEV void front_clear(const FrontFrame &F) {
    if (!F.clear) return;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; 4 * i < F.clear_words;
         i += gridDim.x * blockDim.x) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) v[j] = 4 * i + j == F.clear_idx ? F.clear_val : 0u;
        if (4 * i + 4 <= F.clear_words) {
            *reinterpret_cast<uint4 *>(F.clear + 4 * i) = make_uint4(v[0], v[1], v[2], v[3]);
        } else {
            for (uint32_t j = 0; 4 * i + j < F.clear_words; j++) F.clear[4 * i + j] = v[j];
        }
    }
}
constexpr int FRONT_MAX_FRAMES = 16;
struct FrontTable {
    FrontFrame fr[FRONT_MAX_FRAMES];
    uint32_t n;
};

// Front-end shapes: threshold regions of rx x ry 4x4-pixel tiles, CCL tiles
// of 32 x ccl_rows 2x2-pixel blocks. Results do not depend on them.
struct FrontShapes {
    int thr_rx = 64, thr_ry = 16, ccl_rows = 16;
};
bool front_shapes_supported(const FrontShapes &sh);
// Every FrontShapes value front_shapes_supported accepts.
int front_shape_candidates(FrontShapes *out, int max_out);

// threshold.hip
int launch_decimate(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                    uint8_t *out, int channels = 1);
int launch_threshold(hipStream_t s, const uint8_t *in, size_t w_in, size_t h_in, int f,
                     uint8_t *gray_out, atq_extrema *ext, atq_extrema *fil, uint8_t *thr,
                     int min_diff);
// frames[0..n) (n <= FRONT_MAX_FRAMES), all with shape sh.
int launch_threshold_frames(hipStream_t s, const FrontFrame *frames, int n, const FrontShapes &sh);

// ccl.hip
int launch_ccl(hipStream_t s, const uint8_t *thr, size_t w, size_t h, uint32_t *parent,
               uint32_t *labels, uint32_t *sizes, uint32_t size_cap = 0xFFFFFFFFu);
int launch_ccl_frames(hipStream_t s, const FrontFrame *frames, int n, const FrontShapes &sh);
// Fixed-mode labelling (white 8-connected, black 4-connected, labels root + 1,
// sizes of w * h + 1 entries with sizes[0] = 0).
int launch_ccl_fixed_frames(hipStream_t s, const FrontFrame *frames, int n);

// boundary.hip
int launch_find_boundaries_sparse(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                  size_t w, size_t h, atq_boundary_point *points,
                                  uint64_t *keys, int min_size);
// The detector's open-addressing table of cluster pairs: per slot the pair
// key (stored as ~key, 0 = empty), its record count and its ClusterBounds
// reduction (bbox as 0xFFFF - x_min, x_max, 0xFFFF - y_min, y_max, all
// max-reduced; the two gradient sums; the moment sum), so that every field
// starts at zero (slots = mask + 1, a power of two; zeroed before each
// frame).
struct PairTable {
    uint64_t *keys;
    uint32_t *cnt;
    uint32_t *bbox;  // 4 per slot
    unsigned long long *grad;  // gy * 2^32 + gx per slot
    unsigned long long *moment;
    uint32_t mask;
};
// Buffers of the detector's boundary pass: per table slot the verdict /
// rank of its pair, the listed valid pairs (key min << 32 | max, slot,
// record count), rank -> packed key (cap pairs at most, a multiple of 64),
// the ranked pairs' lookup table (vmask + 1 >= 2 * cap words, a power of
// two: (rank << 2lb) | packed key, all ones when empty), the per-tile
// record regions (4 * w * h entries) and the tiles' counts and offsets (2
// per 1024-pixel tile).
struct PairRanks {
    uint32_t *slot_rank;
    uint64_t *vkeys;
    uint32_t *vslots;
    uint32_t *vcnt;
    uint64_t *rank_keys;
    uint32_t cap;
    uint64_t *vtab;
    uint32_t vmask;
    uint2 *sums;        // per 4-pixel thread of k_pair_count (bnd_summary_entries)
    uint32_t *wg_map;   // k_pair_count's LDS slot -> pair-table slot (bnd_map_words)
    uint64_t *tile_keys;
    atq_boundary_point *tile_points;
    uint32_t *tile_counts;
};
struct RsDyn;  // sort.hpp
// Records of the pairs that pass ValidBlobFilter, plane by plane, each plane
// in raster order, keyed for rank_sort (rank keys, or packed keys
// ((min << lb | max) << 2 | plane) when counts->sort_passes is 0), written
// where rank_sort reads them;
// counts->n_boundary, n_valid_pairs and sort_passes set; in rank mode also
// the clusters' bounds and extents (out.cl) and counts n_pairs, n_clusters
// and n_cluster_points. sc must be the
// look-back form (two scans). fx: fixed mode (column 0 local, both
// half-pixel bits decoded; the labels and sizes must come from the
// fixed-mode CCL).
int launch_find_boundaries_filtered(hipStream_t s, const uint32_t *labels, const uint32_t *sizes,
                                    size_t w, size_t h, int min_size, uint32_t label_bits,
                                    const PairTable &table, atq_counts *counts,
                                    const PairRanks &rk, const RsDyn &out, ScanCtx &sc,
                                    bool fx = false);
constexpr unsigned BOUNDARY_TILE_PX = 1024;  // pixels per compaction tile
// Sizes of PairRanks::sums and wg_map for frames of up to max_w x max_h.
size_t bnd_summary_entries(size_t max_w, size_t max_h);
size_t bnd_map_words(size_t max_w, size_t max_h);

// diff.hip
// decode.hip
int launch_decode_payloads(hipStream_t s, const uint8_t *gray, size_t w, size_t h,
                           const atq_quad_corners *quads, size_t n, const uint32_t *n_dev,
                           const atq_tag_layout &layout, atq_tag_code *out);
int launch_match_codes(hipStream_t s, const atq_tag_code *codes, size_t n, const uint32_t *n_dev,
                       const uint64_t *table, size_t n_table, const atq_tag_layout &layout,
                       int max_hamming, atq_tag_match *out);

int launch_diff_forward(hipStream_t s, const float *img, size_t w, size_t h,
                        const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                        size_t nq, float *corners, float *w_out);
int launch_diff_backward(hipStream_t s, const float *img, size_t w, size_t h,
                         const atq_cluster_point *pts, const float *sw, const atq_diff_quad *quads,
                         size_t nq, const float *gcorners, float *gw, float *gimg);

}  // namespace atq
