// pipeline.hpp - buffer layouts of the glue workspace and the detector.
#pragma once
#include "clusters.hpp"
#include "kernels.hpp"
#include "scan.hpp"
#include "segsort.hpp"

namespace atq {

// Capacities for a frame of npx decimated pixels: at most 4 boundary records
// per pixel (one per tested direction) and, since a valid cluster has at
// least 24 points, at most cap/24 + 1 clusters.
// The detector's pair table has pair_slots (a power of two in [2^14, 2^20],
// about one per 64 pixels) slots.
struct Caps {
    uint32_t npx, cap, cap_cl, pair_slots;
    static Caps for_pixels(size_t npx) {
        Caps c;
        c.npx = (uint32_t)npx;
        c.cap = (uint32_t)(4 * npx);
        c.cap_cl = c.cap / 24 + 1;
        c.pair_slots = 1u << 14;
        while (c.pair_slots < (1u << 20) && c.pair_slots < npx / 64) c.pair_slots <<= 1;
        return c;
    }
};

// Temporaries shared by the glue entry points and the detector.
struct Workspace {
    uint64_t *keys_alt;
    uint32_t *vals_alt;
    uint32_t *seg_id, *seg_start, *valid, *seg_cluster;
    atq_cluster_bounds *seg_bounds;
    atq_cluster_point *cpts_tmp;
    int64_t *fit_sums;  // per (cluster, sub-chunk) line-fit sums of long clusters
    uint32_t *long_list;  // long segments of the angle and corner sorts (2 lists)
    uint32_t *tiles;
    atq_quad_corners *qtmp;
    uint32_t *scan;
    uint32_t *hist, *digit_totals;
    atq_counts *counts;

    // Two key buffers of cap entries each for the angle and corner sorts;
    // they reuse seg_bounds (cap entries of 32 bytes), which is dead once the
    // points are assigned to clusters.
    uint64_t *angle_keys() const { return reinterpret_cast<uint64_t *>(seg_bounds); }

    // Long-segment list k (0: angle sort, 1: corner sort) counted at n.
    SegLong long_segs(int k, uint32_t *n, const Caps &c) const {
        const uint32_t cap = seg_long_cap(c.cap);
        return SegLong{n, long_list + (size_t)k * cap, cap};
    }

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
    uint32_t *slot_rank;   // per pair-table slot: verdict / rank of its pair
    uint64_t *vkeys;       // the frame's valid pairs: key and slot
    uint32_t *vslots, *vcnt;
    uint64_t *rank_keys;   // rank -> packed pair key
    uint64_t *vtab;        // ranked pairs' lookup table (4 * rank_cap words)
    uint2 *bnd_sums;       // the pair count's per-thread record summaries
    uint32_t *bnd_map;     // the pair count's LDS slot -> pair-table slot maps
    float2 *centres;       // rank -> the cluster's (cx, cy)
    uint64_t *scan_flags;  // two look-back flag arrays, zeroed before each frame
    uint64_t *pair_keys;   // pair table, zeroed with counts before each frame
    uint32_t *pair_cnt, *pair_bbox;
    unsigned long long *pair_grad;
    unsigned long long *pair_moment;
    uint32_t *long_n;      // long-segment counts of the two segment sorts and the
                           // full-key sort's barrier word, zeroed likewise
    atq_counts *counts;
    atq::Workspace ws;
    atq::FrontShapes shapes;  // threshold region and CCL tile shapes
    size_t w, h;
    int channels;  // of the last frame's input: 1 grey, 3 RGB, 4 RGBA
    bool sorted_in_alt;
    bool timing;
    // a batch's frame: several frames share the device, so the long
    // clusters' angle sort keeps the work-efficient LDS merge (one read and
    // write of each cluster) over the latency-saving global passes
    bool in_batch = false;
    hipEvent_t ev[ATQ_N_STAGES + 1];
};
