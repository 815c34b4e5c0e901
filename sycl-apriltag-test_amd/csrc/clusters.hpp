// clusters.hpp - launchers of the cluster and fit_quad stages.
#pragma once
#include "records.hpp"
#include "segsort.hpp"
#include "sort.hpp"

namespace atq {

struct ScanCtx;  // scan.hpp

// clusters.hip
int launch_segments(hipStream_t s, const uint64_t *keys, uint32_t key_shift, const uint32_t *n_dev,
                    atq_counts *counts, uint32_t cap, uint32_t *seg_id, uint32_t *seg_start,
                    ScanCtx &sc);
int launch_bounds_filter(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_start,
                         atq_counts *counts, uint32_t cap, uint32_t w, uint32_t h, uint32_t *valid,
                         atq_cluster_bounds *seg_bounds, uint32_t *seg_cluster,
                         atq_cluster_bounds *bounds, ScanCtx &sc, bool fx = false);
int launch_extents(hipStream_t s, const atq_cluster_bounds *bounds, atq_counts *counts, uint32_t cap,
                   atq_cluster_extents *ext, ScanCtx &sc);
int launch_assign(hipStream_t s, const atq_boundary_point *pts, const uint32_t *seg_id,
                  const uint32_t *seg_start, const uint32_t *seg_cluster,
                  const atq_cluster_bounds *bounds, const atq_cluster_extents *ext,
                  const uint32_t *n_dev, uint32_t cap, atq_cluster_point *cpts, uint32_t *cidx,
                  bool fx = false);
// Angle sort of every cluster's points (src/main.cpp:737-750), the points
// in that order to `out` and their line-fit prefix sums to `lf` (may be
// null; src/main.cpp:768-801). max_cluster bounds the cluster length.
// keys0/keys1: two key buffers of cap entries; slot_sums: 6 * (cap_cl + cap
// / FIT_SUB + 2) int64; lng: a long-segment list of seg_long_cap(cap)
// entries whose count is zero. fx: fixed mode (record decode, weight row).
// cidx (may be null) receives every point's cluster index. wts: cap
// entries of scratch for the points' line-fit weights (required with lf).
// throughput: the frame shares the device with others (a batch), so the
// long clusters merge in LDS whatever their bound (seg_lds_merge).
constexpr uint32_t FIT_SUB = 1024;  // the line fit's sub-chunks (a quarter of SEG_CHUNK)
int launch_angle_fit(hipStream_t s, const atq_cluster_point *in, atq_cluster_point *out,
                     const atq_cluster_extents *ext, const atq_counts *counts, uint32_t cap_cl,
                     uint32_t max_cluster, const uint8_t *grey, uint32_t w, uint32_t h,
                     atq_line_fit_point *lf, uint64_t *keys0, uint64_t *keys1, int64_t *slot_sums,
                     SegLong lng, bool fx = false, uint32_t *cidx = nullptr,
                     uint16_t *wts = nullptr, bool throughput = false);
// Line-fit prefix sums of points already in angle order (one work-group per
// cluster).
int launch_cluster_fit(hipStream_t s, const atq_cluster_point *in, const atq_cluster_extents *ext,
                       const uint32_t *n_clusters_dev, uint32_t cap_cl, const uint8_t *grey,
                       uint32_t w, uint32_t h, atq_line_fit_point *lf);
int launch_assign_search(hipStream_t s, const atq_boundary_point *pts, const uint32_t *n_dev,
                         uint32_t cap, const atq_cluster_bounds *bounds, const atq_counts *counts,
                         const atq_cluster_extents *ext, atq_cluster_point *cpts, uint32_t *cidx);
int launch_run_extents(hipStream_t s, const uint32_t *cidx, const uint32_t *n_dev, uint32_t cap,
                       atq_cluster_extents *ext, uint32_t *n_runs_dev, ScanCtx &sc);

// fitquad.hip
// fit_lines (src/fit_quad.cpp:153-375, per point). errs (cap
// floats, may be null) holds the window errors between two kernels. fx:
// fixed mode (global index 0 is a candidate too).
int launch_fit_lines(hipStream_t s, const atq_line_fit_point *lf, const uint32_t *cidx,
                     const atq_cluster_extents *ext, const uint32_t *n_points_dev, uint32_t cap,
                     atq_corner *sparse, uint32_t *ccount, float *errs = nullptr,
                     bool fx = false);
int launch_set_u32(hipStream_t s, uint32_t *p, uint32_t v);
// Zeroes p[0 .. min(*n_dev, cap)) (n_dev null: cap words), except that
// p[set_idx] gets set_val.
int launch_zero_u32(hipStream_t s, uint32_t *p, const uint32_t *n_dev, uint32_t cap,
                    uint32_t set_idx = 0xFFFFFFFFu, uint32_t set_val = 0);
// Corners compacted and grouped into per-cluster peak segments, each in
// (error descending, index) order; scratch: 2 * cap u64; lng as for
// launch_angle_fit. top_only: only each segment's min(10, n) first peaks in
// that order (all that do_indexing reads), the others after them in the
// compaction's order but for the ones those ten displaced.
int launch_corner_segments(hipStream_t s, const atq_corner *sparse, const uint32_t *n_sparse_dev,
                           uint32_t cap, atq_counts *counts, uint32_t cap_cl, uint32_t max_seg,
                           atq_peak_extents *peaks, atq_corner *compacted, uint64_t *scratch,
                           SegLong lng, ScanCtx &sc, bool top_only = false);
int launch_do_indexing(hipStream_t s, const atq_peak_extents *peaks, const uint32_t *n_seg_dev,
                       uint32_t cap, const atq_corner *corners, const atq_line_fit_point *lf,
                       const uint32_t *n_points_dev, uint32_t n_points_host,
                       const atq_cluster_extents *ext, atq_fitted_quad *fitted,
                       uint32_t *oob_counter, bool fx = false);
// process_fitted_quads (minus decoding) and the passing quads' compaction in
// one look-back scan pass; tmp: cap entries of scratch.
int launch_quads(hipStream_t s, const atq_fitted_quad *fitted, const uint32_t *n_fitted_dev,
                 uint32_t cap, atq_quad_corners *tmp, atq_quad_corners *out,
                 uint32_t *n_out_dev, ScanCtx &sc, int min_tag_width);

}  // namespace atq
