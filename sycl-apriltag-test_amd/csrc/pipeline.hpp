// pipeline.hpp - buffer layouts of the glue workspace and the detector.
#pragma once
#include "clusters.hpp"
#include "kernels.hpp"

namespace atq {

// Capacities for a frame of npx decimated pixels: at most 4 boundary records
// per pixel (one per tested direction) and, since a valid cluster has at
// least 24 points, at most cap/24 + 1 clusters.
struct Caps {
    uint32_t npx, cap, cap_cl;
    static Caps for_pixels(size_t npx) {
        Caps c;
        c.npx = (uint32_t)npx;
        c.cap = (uint32_t)(4 * npx);
        c.cap_cl = c.cap / 24 + 1;
        return c;
    }
};

// Temporaries shared by the glue entry points and the detector.
struct Workspace {
    uint64_t *keys_alt;
    uint32_t *vals_alt;
    uint32_t *seg_id, *seg_start, *valid, *seg_cluster;
    atq_cluster_bounds *seg_bounds;
    uint64_t *scratch64;
    atq_cluster_point *cpts_tmp;
    uint32_t *tiles;
    atq_quad_corners *qtmp;
    uint32_t *qflags;
    uint32_t *scan;
    uint32_t *hist, *digit_base;
    atq_counts *counts;

    // The angle sort's two key buffers (2 * cap entries) reuse seg_bounds
    // (cap entries of 32 bytes), which is dead once points are assigned.
    uint64_t *angle_keys() const { return reinterpret_cast<uint64_t *>(seg_bounds); }

    // Carves the layout from `base` (nullptr: only measures). Returns bytes.
    size_t carve(char *base, const Caps &c);
};

}  // namespace atq

struct atq_detector {
    size_t max_w_in, max_h_in, max_w, max_h;
    atq_params params;
    atq::Caps caps;
    char *arena;  // one device allocation for every buffer below
    uint8_t *gray, *thr;
    atq_extrema *ext, *fil;
    uint32_t *parent, *labels, *sizes;
    uint64_t *keys;
    uint32_t *vals;
    atq_cluster_bounds *bounds;
    atq_cluster_extents *cext;
    atq_cluster_point *cpts;
    uint32_t *cidx;
    atq_line_fit_point *lf;
    atq_corner *sparse, *corners;
    atq_peak_extents *peaks;
    atq_fitted_quad *fitted;
    atq_quad_corners *quads;
    atq_counts *counts;
    atq::Workspace ws;
    size_t w, h;
    bool sorted_in_alt;
    bool timing;
    hipEvent_t ev[ATQ_N_STAGES + 1];
};
