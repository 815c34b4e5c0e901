// pipeline.cpp - the whole-frame detector and the glue entry points.
//
// The detector runs threshold -> quads as one stream-ordered chain of
// kernels. Every count (boundary records, segments, clusters, points,
// corners, quads) stays in device memory and later kernels read it there,
// so the host never waits between stages (the reference synchronises after
// every oneDPL call that returns a count, src/main.cpp:470-907). Buffers are
// sized once, at creation, for the largest frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <tuple>
#include <vector>

#include "pipeline.hpp"

using namespace atq;

namespace atq {

namespace {

template <typename T>
T *carve_one(char *base, size_t &off, size_t count) {
    off = (off + 255) & ~(size_t)255;
    T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
    off += sizeof(T) * (count ? count : 1);
    return p;
}

struct SparseFlag {
    const uint32_t *pts;
    ATQ_DEV uint32_t operator()(uint32_t i) const { return pts[i] != 0u; }
};
struct SparseOut {
    const uint32_t *pts;
    const uint64_t *keys;
    uint32_t *out_pts;
    uint64_t *out_keys;
    ATQ_DEV void operator()(uint32_t i, uint32_t pre, uint32_t v) const {
        if (v) {
            out_pts[pre] = pts[i];
            out_keys[pre] = keys[i];
        }
    }
};

}  // namespace

size_t Workspace::carve(char *base, const Caps &c) {
    size_t off = 0;
    keys_alt = carve_one<uint64_t>(base, off, c.cap);
    vals_alt = carve_one<uint32_t>(base, off, c.cap);
    seg_id = carve_one<uint32_t>(base, off, c.cap);
    seg_start = carve_one<uint32_t>(base, off, (size_t)c.cap + 1);
    valid = carve_one<uint32_t>(base, off, c.cap);
    seg_cluster = carve_one<uint32_t>(base, off, c.cap);
    seg_bounds = carve_one<atq_cluster_bounds>(base, off, c.cap);
    cpts_tmp = carve_one<atq_cluster_point>(base, off, c.cap);
    fit_sums = carve_one<int64_t>(base, off, 6 * ((size_t)c.cap_cl + c.cap / FIT_SUB + 2));
    long_list = carve_one<uint32_t>(base, off, 2 * (size_t)seg_long_cap(c.cap));
    tiles = carve_one<uint32_t>(base, off, 8 * ((size_t)c.npx / BOUNDARY_TILE_PX + 1));
    qtmp = carve_one<atq_quad_corners>(base, off, c.cap_cl);
    scan = carve_one<uint32_t>(base, off, SCAN_GROUPS + 1);
    hist = carve_one<uint32_t>(base, off, (size_t)RS_MAX_DIGITS * RS_GROUPS);
    digit_totals = carve_one<uint32_t>(base, off, RS_MAX_DIGITS);
    counts = carve_one<atq_counts>(base, off, 1);
    return (off + 255) & ~(size_t)255;
}

}  // namespace atq

namespace {

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

int check_dims(size_t w, size_t h) {
    if (w < 8 || h < 8 || w > 8192 || h > 8192 || (w & 3) || (h & 3)) return ATQ_ERR_BAD_DIMS;
    return ATQ_OK;
}

size_t decimated(size_t n, int f) { return f > 1 ? 1 + (n - 1) / f : n; }

// The largest frame (in pixels) a workspace of `bytes` serves.
Caps caps_from_workspace(size_t bytes) {
    Workspace ws;
    size_t lo = 8, hi = (size_t)8192 * 8192;
    while (lo < hi) {
        size_t mid = lo + (hi - lo + 1) / 2;
        if (ws.carve(nullptr, Caps::for_pixels(mid)) <= bytes) lo = mid;
        else hi = mid - 1;
    }
    return Caps::for_pixels(lo);
}

int take_workspace(Workspace &ws, void *workspace, size_t bytes, size_t w, size_t h) {
    if (!workspace) return ATQ_ERR_BAD_ARG;
    Caps c = Caps::for_pixels(w * h);
    if (ws.carve(nullptr, c) > bytes) return ATQ_ERR_BAD_ARG;
    ws.carve(static_cast<char *>(workspace), c);
    return ATQ_OK;
}

// Look-back flags per scan array: enough tiles for the longest scan (at
// most cap elements; scans of at most LB_SMALL_CAP elements take tiles of
// LB_TILE / 2) and for the pair table's slots (512 per tile).
uint32_t scan_flag_count(const Caps &c) {
    return std::max({c.cap / LB_TILE, std::min(c.cap, LB_SMALL_CAP) / (LB_TILE / 2),
                     c.pair_slots / 512}) + 2;
}

// Valid cluster pairs the detector ranks (more: the full-key sort runs).
uint32_t rank_cap(const Caps &c) { return std::max(64u, std::min(16384u, c.pair_slots / 4)); }

// Detector buffers, carved from one arena.
size_t carve_detector(atq_detector *d, char *base) {
    const Caps &c = d->caps;
    size_t off = 0;
    d->gray = carve_one<uint8_t>(base, off, c.npx);
    d->thr = carve_one<uint8_t>(base, off, c.npx);
    d->ext = carve_one<atq_extrema>(base, off, c.npx / 16 + 1);
    d->fil = carve_one<atq_extrema>(base, off, c.npx / 16 + 1);
    d->parent = carve_one<uint32_t>(base, off, c.npx);
    d->labels = carve_one<uint32_t>(base, off, c.npx);
    d->sizes = carve_one<uint32_t>(base, off, c.npx + 1);  // fixed mode: root + 1
    d->keys = carve_one<uint64_t>(base, off, c.cap);
    d->vals = carve_one<uint32_t>(base, off, c.cap);
    d->bounds = carve_one<atq_cluster_bounds>(base, off, c.cap_cl);
    d->cext = carve_one<atq_cluster_extents>(base, off, c.cap_cl);
    d->cpts = carve_one<atq_cluster_point>(base, off, c.cap);
    d->cidx = carve_one<uint32_t>(base, off, c.cap);
    d->lf = carve_one<atq_line_fit_point>(base, off, c.cap);
    d->sparse = carve_one<atq_corner>(base, off, c.cap);
    d->corners = carve_one<atq_corner>(base, off, c.cap);
    d->peaks = carve_one<atq_peak_extents>(base, off, c.cap_cl);
    d->fitted = carve_one<atq_fitted_quad>(base, off, c.cap_cl);
    d->quads = carve_one<atq_quad_corners>(base, off, c.cap_cl);
    d->slot_rank = carve_one<uint32_t>(base, off, c.pair_slots);
    d->vkeys = carve_one<uint64_t>(base, off, rank_cap(c));
    d->vslots = carve_one<uint32_t>(base, off, rank_cap(c));
    d->vcnt = carve_one<uint32_t>(base, off, rank_cap(c));
    d->centres = carve_one<float2>(base, off, rank_cap(c));
    d->rank_keys = carve_one<uint64_t>(base, off, rank_cap(c));
    d->vtab = carve_one<uint64_t>(base, off, 4 * (size_t)rank_cap(c));
    d->bnd_sums = carve_one<uint2>(base, off, bnd_summary_entries(d->max_w, d->max_h));
    d->bnd_map = carve_one<uint32_t>(base, off, bnd_map_words(d->max_w, d->max_h));
    d->scan_flags = carve_one<uint64_t>(base, off, 2 * (size_t)scan_flag_count(c));
    d->pair_keys = carve_one<uint64_t>(base, off, c.pair_slots);
    d->pair_cnt = carve_one<uint32_t>(base, off, c.pair_slots);
    d->pair_bbox = carve_one<uint32_t>(base, off, 4 * (size_t)c.pair_slots);
    d->pair_grad = carve_one<unsigned long long>(base, off, (size_t)c.pair_slots);
    d->pair_moment = carve_one<unsigned long long>(base, off, c.pair_slots);
    d->long_n = carve_one<uint32_t>(base, off, 4);  // [2]: the full-key sort's barrier
    d->counts = carve_one<atq_counts>(base, off, 1);
    off = (off + 255) & ~(size_t)255;
    size_t ws_bytes = d->ws.carve(base ? base + off : nullptr, c);
    return off + ws_bytes;
}

#define TRY(x)                      \
    do {                            \
        int r_ = (x);               \
        if (r_ != ATQ_OK) return r_; \
    } while (0)

// Validates a frame's dims against the detector and records them.
int frame_setup(atq_detector *d, size_t w_in, size_t h_in, int channels = 1) {
    if (channels != 1 && channels != 3 && channels != 4) return ATQ_ERR_BAD_ARG;
    d->channels = channels;
    const int f = d->params.decimate;
    const size_t w = decimated(w_in, f), h = decimated(h_in, f);
    TRY(check_dims(w, h));
    if (w_in > d->max_w_in || h_in > d->max_h_in || w * h > d->caps.npx) return ATQ_ERR_CAPACITY;
    d->w = w;
    d->h = h;
    return ATQ_OK;
}

// The front-end description of the detector's frame (after frame_setup).
FrontFrame front_frame(const atq_detector *d, const uint8_t *gray_in, size_t w_in, size_t h_in) {
    FrontFrame fr{};
    fr.in = gray_in;
    fr.gray = d->gray;
    // no later stage reads the tile extrema: the detector does not write them
    fr.ext = nullptr;
    fr.fil = nullptr;
    fr.thr = d->thr;
    fr.parent = d->parent;
    fr.labels = d->labels;
    fr.sizes = d->sizes;
    fr.w_in = (uint32_t)w_in;
    fr.h_in = (uint32_t)h_in;
    fr.f = (uint32_t)d->params.decimate;
    fr.w = (uint32_t)d->w;
    fr.h = (uint32_t)d->h;
    // sizes only decide "at least min_component_size" here
    fr.size_cap = (uint32_t)std::max(d->params.min_component_size, 1);
    fr.min_diff = d->params.min_white_black_diff;
    fr.channels = (uint32_t)d->channels;
    // one fill for the later stages: the scan flags, the pair table, the
    // long-segment counts and the counts (carved contiguously), with the
    // counts' key width set
    uint32_t *zero_begin = reinterpret_cast<uint32_t *>(d->scan_flags);
    fr.clear = zero_begin;
    fr.clear_words = (uint32_t)((reinterpret_cast<char *>(d->counts + 1) -
                                 reinterpret_cast<char *>(zero_begin)) / 4);
    fr.clear_idx = (uint32_t)(reinterpret_cast<uint32_t *>(&d->counts->key_label_bits) - zero_begin);
    fr.clear_val = (uint32_t)radix_plan(d->w * d->h, 2).label_bits;
    return fr;
}

int mark(atq_detector *d, hipStream_t s, int i) {
    if (d->timing && hipEventRecord(d->ev[i], s) != hipSuccess) return ATQ_ERR_HIP;
    return ATQ_OK;
}

// boundary -> quads on `s`, after the frame's threshold and CCL.
int run_back(atq_detector *d, hipStream_t s, const uint8_t *gray_in) {
    const int f = d->params.decimate;
    const size_t w = d->w, h = d->h;
    const Caps c = Caps::for_pixels(w * h);  // this frame's own bounds
    Workspace &ws = d->ws;
    atq_counts *cnt = d->counts;
    ScanCtx sc;
    sc.flags = d->scan_flags;
    sc.flag_count = scan_flag_count(d->caps);
    sc.error = &cnt->overflow;
    // the scan flags, the pair table, the long-segment counts and the counts
    // (with the key width set) were zeroed by this frame's last CCL kernel
    const RadixPlan plan = radix_plan(w * h, 2);
    // the grey image the line fit samples: the input itself only when it is
    // grey and undecimated
    const uint8_t *grey = f > 1 || d->channels != 1 ? d->gray : gray_in;
    auto mark = [&](int i) { return ::mark(d, s, i); };
    // Boundary records of the pairs that pass ValidBlobFilter, in the
    // reference's order, keyed by their pair's rank among the valid pairs
    // (or, in the fallback, the packed pair key min<<lb|max) with the plane
    // in 2 low bits; the stable sort then yields the reference's
    // (pair, plane, y, x) order, and its last pass restores the packed keys.
    const PairTable table{d->pair_keys, d->pair_cnt,    d->pair_bbox,
                          d->pair_grad, d->pair_moment, d->caps.pair_slots - 1};
    const bool fx = d->params.mode == ATQ_MODE_FIXED;
    RsClusters cl{};
    cl.bounds = d->bounds;
    cl.ext = d->cext;
    cl.centre = d->centres;
    cl.cpts = ws.cpts_tmp;
    cl.cidx = d->cidx;
    cl.seg_id = ws.seg_id;
    cl.seg_start = ws.seg_start;
    cl.valid = ws.valid;
    cl.seg_cluster = ws.seg_cluster;
    cl.seg_bounds = ws.seg_bounds;
    cl.part = ws.scan;
    cl.counts = cnt;
    cl.w = (uint32_t)w;
    cl.h = (uint32_t)h;
    cl.fx = fx ? 1u : 0u;
    const RsDyn io{d->keys,        d->vals,          ws.keys_alt, ws.vals_alt,
                   &cnt->sort_passes, &cnt->rank_shift, d->rank_keys, cl};
    // the per-tile record regions live in the cluster-bounds area, which is
    // free until the fallback's cluster stage (12 of its 32 bytes per entry)
    uint64_t *tile_keys = reinterpret_cast<uint64_t *>(ws.seg_bounds);
    const PairRanks rk{d->slot_rank, d->vkeys, d->vslots, d->vcnt, d->rank_keys, rank_cap(d->caps),
                       d->vtab, 4 * rank_cap(d->caps) - 1, d->bnd_sums, d->bnd_map, tile_keys, reinterpret_cast<atq_boundary_point *>(tile_keys + c.cap),
                       ws.tiles};
    TRY(launch_find_boundaries_filtered(s, d->labels, d->sizes, w, h, d->params.min_component_size,
                                        (uint32_t)plan.label_bits, table, cnt, rk, io, sc, fx));
    TRY(mark(1 + ATQ_STAGE_BOUNDARY));
    // the sorted records end in the alt buffers, the clusters' points in
    // cpts_tmp with their bounds and extents (the angle sort records each
    // point's cluster index)
    TRY(rank_sort(s, 2 * (uint32_t)plan.label_bits + 2, io, &cnt->n_boundary, c.cap, ws.hist,
                  ws.digit_totals, &d->long_n[2], &cnt->overflow));
    TRY(mark(1 + ATQ_STAGE_KEY_SORT));
    d->sorted_in_alt = true;
    TRY(mark(1 + ATQ_STAGE_CLUSTERS));
    // a valid cluster holds at most 2 * (2w + 2h) points (valid_blob)
    const uint32_t max_cluster = (uint32_t)std::min<size_t>(c.cap, 4 * (w + h));
    uint64_t *ak = ws.angle_keys();
    // the line-fit weights use ws.valid, free from the key sort until fit_lines
    TRY(launch_angle_fit(s, ws.cpts_tmp, d->cpts, d->cext, cnt, c.cap_cl, max_cluster, grey,
                         (uint32_t)w, (uint32_t)h, d->lf, ak, ak + c.cap, ws.fit_sums,
                         ws.long_segs(0, &d->long_n[0], c), fx, d->cidx,
                         reinterpret_cast<uint16_t *>(ws.valid), d->in_batch));
    TRY(mark(1 + ATQ_STAGE_ANGLE_FIT));
    TRY(launch_fit_lines(s, d->lf, d->cidx, d->cext, &cnt->n_cluster_points, c.cap, d->sparse,
                         nullptr, reinterpret_cast<float *>(ws.valid), fx));
    TRY(mark(1 + ATQ_STAGE_FIT_LINES));
    TRY(launch_corner_segments(s, d->sparse, &cnt->n_cluster_points, c.cap, cnt, c.cap_cl,
                               max_cluster, d->peaks, d->corners, ws.angle_keys(),
                               ws.long_segs(1, &d->long_n[1], c), sc, true));
    TRY(mark(1 + ATQ_STAGE_CORNERS));
    TRY(launch_do_indexing(s, d->peaks, &cnt->n_segments, c.cap_cl, d->corners, d->lf,
                           &cnt->n_cluster_points, 0, d->cext, d->fitted, &cnt->oob_reads, fx));
    TRY(mark(1 + ATQ_STAGE_INDEXING));
    TRY(launch_quads(s, d->fitted, &cnt->n_segments, c.cap_cl, ws.qtmp, d->quads,
                     &cnt->n_quads, sc, d->params.min_tag_width));
    TRY(mark(1 + ATQ_STAGE_QUADS));
    return ATQ_OK;
}

// threshold -> quads, all on `s`.
int run_pipeline(atq_detector *d, hipStream_t s, const uint8_t *gray_in, size_t w_in, size_t h_in,
                 int channels = 1) {
    TRY(frame_setup(d, w_in, h_in, channels));
    TRY(mark(d, s, 0));
    const FrontFrame fr = front_frame(d, gray_in, w_in, h_in);
    TRY(launch_threshold_frames(s, &fr, 1, d->shapes));
    TRY(mark(d, s, 1 + ATQ_STAGE_THRESHOLD));
    if (d->params.mode == ATQ_MODE_FIXED) TRY(launch_ccl_fixed_frames(s, &fr, 1));
    else TRY(launch_ccl_frames(s, &fr, 1, d->shapes));
    TRY(mark(d, s, 1 + ATQ_STAGE_CCL));
    return run_back(d, s, gray_in);
}

}  // namespace

extern "C" {

size_t atq_glue_workspace_size(size_t width, size_t height) {
    Workspace ws;
    return ws.carve(nullptr, Caps::for_pixels(width * height));
}

int atq_group_boundaries(void *stream, const atq_boundary_point *points, const uint64_t *blob_labels,
                         size_t width, size_t height, atq_boundary_point *out_points,
                         uint64_t *out_keys, uint32_t *count_dev, void *workspace,
                         size_t workspace_bytes) {
    if (!points || !blob_labels || !out_points || !out_keys || !count_dev) return ATQ_ERR_BAD_ARG;
    TRY(check_dims(width, height));
    Workspace ws;
    TRY(take_workspace(ws, workspace, workspace_bytes, width, height));
    ScanCtx sc;
    sc.scratch = ws.scan;
    hipStream_t s = S(stream);
    const uint32_t n = (uint32_t)(4 * width * height);
    // the element count of the sparse planes, as a device scalar
    uint32_t *n_dev = &ws.counts->reserved[0];
    TRY(launch_set_u32(s, n_dev, n));
    const uint32_t *pts = reinterpret_cast<const uint32_t *>(points);
    uint32_t *opts = reinterpret_cast<uint32_t *>(out_points);
    TRY(device_scan(s, n_dev, n, SparseFlag{pts}, SparseOut{pts, blob_labels, opts, out_keys},
                    sc, count_dev));
    const RadixPlan plan = radix_plan(width * height, 0);
    const uint32_t lb = (uint32_t)plan.label_bits;
    TRY(launch_pack_keys(s, out_keys, out_keys, count_dev, n, lb, false));
    bool alt = false;
    TRY(radix_sort_packed(s, plan, out_keys, opts, ws.keys_alt, ws.vals_alt, count_dev, n, ws.hist,
                          ws.digit_totals, &alt));
    // unpack into the caller's buffer (from wherever the last pass wrote)
    TRY(launch_pack_keys(s, alt ? ws.keys_alt : out_keys, out_keys, count_dev, n, lb, true));
    if (alt)
        ATQ_HIP_CHECK(hipMemcpyAsync(opts, ws.vals_alt, sizeof(uint32_t) * n,
                                     hipMemcpyDeviceToDevice, s));
    // Stream-ordered copy of the count back is not needed: it is on device.
    return ATQ_OK;
}

int atq_cluster_bounds_filter(void *stream, const atq_boundary_point *points, const uint64_t *keys,
                              const uint32_t *n_points_dev, size_t width, size_t height,
                              atq_cluster_bounds *out_bounds, atq_counts *counts_dev,
                              void *workspace, size_t workspace_bytes) {
    if (!points || !keys || !n_points_dev || !out_bounds || !counts_dev) return ATQ_ERR_BAD_ARG;
    TRY(check_dims(width, height));
    Workspace ws;
    TRY(take_workspace(ws, workspace, workspace_bytes, width, height));
    ScanCtx sc;
    sc.scratch = ws.scan;
    hipStream_t s = S(stream);
    const Caps c = Caps::for_pixels(width * height);
    TRY(launch_segments(s, keys, 0, n_points_dev, counts_dev, c.cap, ws.seg_id, ws.seg_start,
                        sc));
    return launch_bounds_filter(s, points, ws.seg_start, counts_dev, c.cap, (uint32_t)width,
                                (uint32_t)height, ws.valid, ws.seg_bounds, ws.seg_cluster,
                                out_bounds, sc);
}

int atq_cluster_points_sort(void *stream, const atq_boundary_point *points,
                            const uint32_t *n_points_dev, const atq_cluster_bounds *bounds,
                            atq_counts *counts_dev, atq_cluster_point *out_points,
                            uint32_t *out_cluster_index, atq_cluster_extents *out_extents,
                            void *workspace, size_t workspace_bytes) {
    if (!points || !n_points_dev || !bounds || !counts_dev || !out_points || !out_cluster_index ||
        !out_extents || !workspace)
        return ATQ_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Workspace ws;
    const Caps c = caps_from_workspace(workspace_bytes);
    ws.carve(static_cast<char *>(workspace), c);
    ScanCtx sc;
    sc.scratch = ws.scan;
    TRY(launch_extents(s, bounds, counts_dev, c.cap_cl, out_extents, sc));
    TRY(launch_assign_search(s, points, n_points_dev, c.cap, bounds, counts_dev, out_extents,
                             ws.cpts_tmp, out_cluster_index));
    uint64_t *ak = ws.angle_keys();
    // a valid cluster holds at most 2 * (2w + 2h) points (valid_blob), and
    // w + h <= npx / 8 + 8 for frames of at least 8 pixels a side
    const uint32_t max_cluster = std::min<uint32_t>(c.cap, c.npx / 2 + 32);
    uint32_t *n_long = &ws.counts->reserved[1];
    TRY(launch_set_u32(s, n_long, 0));
    return launch_angle_fit(s, ws.cpts_tmp, out_points, out_extents, counts_dev, c.cap_cl,
                            max_cluster, nullptr, 0, 0, nullptr, ak, ak + c.cap, ws.fit_sums,
                            ws.long_segs(0, n_long, c));
}

int atq_linefit_prefix(void *stream, const atq_cluster_point *points,
                       const atq_cluster_extents *extents, const atq_counts *counts_dev,
                       const uint8_t *grayscale, size_t width, size_t height,
                       atq_line_fit_point *out) {
    if (!points || !extents || !counts_dev || !grayscale || !out) return ATQ_ERR_BAD_ARG;
    TRY(check_dims(width, height));
    const Caps c = Caps::for_pixels(width * height);
    return launch_cluster_fit(S(stream), points, extents, &counts_dev->n_clusters, c.cap_cl,
                              grayscale, (uint32_t)width, (uint32_t)height, out);
}

int atq_fit_lines(void *stream, const atq_line_fit_point *points, const uint32_t *cluster_indices,
                  const atq_cluster_extents *cluster_extents, size_t points_size,
                  atq_corner *found_corners) {
    if (!points || !cluster_indices || !cluster_extents || !found_corners) return ATQ_ERR_BAD_ARG;
    if (points_size > 0xFFFFFFFFu) return ATQ_ERR_CAPACITY;
    hipStream_t s = S(stream);
    if (points_size) {
        TRY(launch_fit_lines(s, points, cluster_indices, cluster_extents, nullptr,
                             (uint32_t)points_size, found_corners, nullptr));
    }
    ATQ_HIP_CHECK(hipStreamSynchronize(s));  // blocking, src/fit_quad.cpp:374
    return ATQ_OK;
}

int atq_corner_segments(void *stream, const atq_corner *found_corners, atq_counts *counts_dev,
                        atq_corner *compacted, atq_peak_extents *segments, void *workspace,
                        size_t workspace_bytes) {
    if (!found_corners || !counts_dev || !compacted || !segments || !workspace)
        return ATQ_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Workspace ws;
    const Caps c = caps_from_workspace(workspace_bytes);
    ws.carve(static_cast<char *>(workspace), c);
    ScanCtx sc;
    sc.scratch = ws.scan;
    uint32_t *n_long = &ws.counts->reserved[1];
    TRY(launch_set_u32(s, n_long, 0));
    return launch_corner_segments(s, found_corners, &counts_dev->n_cluster_points, c.cap, counts_dev,
                                  c.cap_cl, std::min<uint32_t>(c.cap, c.npx / 2 + 32), segments,
                                  compacted, ws.angle_keys(), ws.long_segs(1, n_long, c), sc);
}

int atq_do_indexing(void *stream, const atq_peak_extents *extents, size_t extents_count,
                    const atq_corner *compacted_corners, const atq_line_fit_point *points,
                    size_t n_points, const atq_cluster_extents *cluster_extents,
                    atq_fitted_quad *fitted) {
    if (!extents || !compacted_corners || !points || !cluster_extents || !fitted)
        return ATQ_ERR_BAD_ARG;
    if (extents_count > 0xFFFFFFFFu || n_points > 0xFFFFFFFFu) return ATQ_ERR_CAPACITY;
    hipStream_t s = S(stream);
    if (extents_count) {
        TRY(launch_do_indexing(s, extents, nullptr, (uint32_t)extents_count, compacted_corners,
                               points, nullptr, (uint32_t)n_points, cluster_extents, fitted,
                               nullptr));
    }
    ATQ_HIP_CHECK(hipStreamSynchronize(s));  // blocking, src/fit_quad.cpp:534
    return ATQ_OK;
}

int atq_quads_from_fits(void *stream, const atq_fitted_quad *fitted, size_t n_fitted,
                        atq_quad_corners *out, uint32_t *n_quads_dev, void *workspace,
                        size_t workspace_bytes) {
    if (!fitted || !out || !n_quads_dev || !workspace) return ATQ_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    Workspace ws;
    const Caps c = caps_from_workspace(workspace_bytes);
    if (n_fitted > c.cap_cl) return ATQ_ERR_CAPACITY;
    ws.carve(static_cast<char *>(workspace), c);
    ScanCtx sc;
    sc.scratch = ws.scan;
    uint32_t *n_dev = &ws.counts->reserved[1];
    TRY(launch_set_u32(s, n_dev, (uint32_t)n_fitted));
    return launch_quads(s, fitted, n_dev, c.cap_cl, ws.qtmp, out, n_quads_dev, sc,
                        8);
}

int atq_detector_create(size_t max_width_in, size_t max_height_in, const atq_params *params,
                        atq_detector **out) {
    if (!out) return ATQ_ERR_BAD_ARG;
    *out = nullptr;
    atq_params p{1, 5, 25, 8, 0, ATQ_MODE_COMPAT};
    if (params) p = *params;
    if (p.decimate < 1) return ATQ_ERR_BAD_ARG;
    if (p.mode != ATQ_MODE_COMPAT && p.mode != ATQ_MODE_FIXED) return ATQ_ERR_BAD_ARG;
    if (p.pair_table_log2 != 0 && (p.pair_table_log2 < 8 || p.pair_table_log2 > 24))
        return ATQ_ERR_BAD_ARG;
    const size_t w = decimated(max_width_in, p.decimate), h = decimated(max_height_in, p.decimate);
    TRY(check_dims(w, h));
    if (int r = atq_init()) return r;
    atq_detector *d = new (std::nothrow) atq_detector();
    if (!d) return ATQ_ERR_BAD_ARG;
    d->max_w_in = max_width_in;
    d->max_h_in = max_height_in;
    d->max_w = w;
    d->max_h = h;
    d->params = p;
    d->caps = Caps::for_pixels(w * h);
    if (p.pair_table_log2) d->caps.pair_slots = 1u << p.pair_table_log2;
    const size_t bytes = carve_detector(d, nullptr);
    if (hipMalloc(&d->arena, bytes) != hipSuccess) {
        delete d;
        return ATQ_ERR_HIP;
    }
    carve_detector(d, d->arena);
    d->w = w;
    d->h = h;
    *out = d;
    return ATQ_OK;
}

int atq_detector_destroy(atq_detector *det) {
    if (!det) return ATQ_ERR_BAD_ARG;
    if (det->timing)
        for (int i = 0; i <= ATQ_N_STAGES; i++) (void)hipEventDestroy(det->ev[i]);
    if (det->arena) (void)hipFree(det->arena);
    delete det;
    return ATQ_OK;
}

int atq_detect_quads(atq_detector *det, void *stream, const uint8_t *gray_in, size_t width_in,
                     size_t height_in) {
    if (!det || !gray_in) return ATQ_ERR_BAD_ARG;
    return run_pipeline(det, S(stream), gray_in, width_in, height_in);
}

int atq_detect_quads_rgb(atq_detector *det, void *stream, const uint8_t *pixels_in,
                         size_t width_in, size_t height_in, int channels) {
    if (!det || !pixels_in || (channels != 3 && channels != 4)) return ATQ_ERR_BAD_ARG;
    return run_pipeline(det, S(stream), pixels_in, width_in, height_in, channels);
}

int atq_detector_set_shapes(atq_detector *det, const atq_shapes *shapes) {
    if (!det || !shapes) return ATQ_ERR_BAD_ARG;
    const FrontShapes sh{shapes->threshold_tiles_x, shapes->threshold_tiles_y,
                         shapes->ccl_block_rows};
    if (!front_shapes_supported(sh)) return ATQ_ERR_BAD_ARG;
    det->shapes = sh;
    return ATQ_OK;
}

int atq_detector_get_shapes(atq_detector *det, atq_shapes *shapes) {
    if (!det || !shapes) return ATQ_ERR_BAD_ARG;
    *shapes = atq_shapes{det->shapes.thr_rx, det->shapes.thr_ry, det->shapes.ccl_rows};
    return ATQ_OK;
}

}  // extern "C"

namespace {

// Shapes chosen so far, per (width_in, height_in, decimate).
std::mutex g_tune_mutex;
std::map<std::tuple<size_t, size_t, int>, FrontShapes> g_tuned;

// Milliseconds of `reps` back-to-back runs of `run` on s (after one warm-up).
template <typename Run>
int time_runs(hipStream_t s, hipEvent_t e0, hipEvent_t e1, int reps, Run run, float *ms) {
    TRY(run());
    ATQ_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) TRY(run());
    ATQ_HIP_CHECK(hipEventRecord(e1, s));
    ATQ_HIP_CHECK(hipEventSynchronize(e1));
    ATQ_HIP_CHECK(hipEventElapsedTime(ms, e0, e1));
    return ATQ_OK;
}

// Times the threshold shapes (CCL shape fixed), then the CCL shapes on the
// thresholded frame, and keeps the fastest of each. Candidates are timed in
// interleaved trials of REPS runs; a candidate's time is its best trial.
int tune_frame(atq_detector *d, hipStream_t s, const uint8_t *gray_in, size_t w_in, size_t h_in,
               FrontShapes *best) {
    TRY(frame_setup(d, w_in, h_in));
    const FrontFrame fr = front_frame(d, gray_in, w_in, h_in);
    FrontShapes cand[16];
    const int nc = std::min(front_shape_candidates(cand, 16), 16);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess) return ATQ_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return ATQ_ERR_HIP;
    }
    constexpr int REPS = 10, TRIALS = 3;
    int rc = ATQ_OK;
    FrontShapes pick = d->shapes;
    float t_thr[16], t_ccl[16];
    for (int i = 0; i < nc; i++) t_thr[i] = t_ccl[i] = 1e30f;
    for (int trial = 0; trial < TRIALS && rc == ATQ_OK; trial++)
        for (int i = 0; i < nc && rc == ATQ_OK; i++) {
            if (cand[i].ccl_rows != pick.ccl_rows) continue;  // threshold shapes first
            float ms = 0;
            rc = time_runs(s, e0, e1, REPS,
                           [&] { return launch_threshold_frames(s, &fr, 1, cand[i]); }, &ms);
            t_thr[i] = std::min(t_thr[i], ms);
        }
    int bi = -1;
    for (int i = 0; i < nc; i++)
        if (cand[i].ccl_rows == pick.ccl_rows && (bi < 0 || t_thr[i] < t_thr[bi])) bi = i;
    if (bi >= 0) {
        pick.thr_rx = cand[bi].thr_rx;
        pick.thr_ry = cand[bi].thr_ry;
    }
    if (rc == ATQ_OK) rc = launch_threshold_frames(s, &fr, 1, pick);
    for (int trial = 0; trial < TRIALS && rc == ATQ_OK; trial++)
        for (int i = 0; i < nc && rc == ATQ_OK; i++) {
            if (cand[i].thr_rx != pick.thr_rx || cand[i].thr_ry != pick.thr_ry) continue;
            float ms = 0;
            rc = time_runs(s, e0, e1, REPS, [&] { return launch_ccl_frames(s, &fr, 1, cand[i]); },
                           &ms);
            t_ccl[i] = std::min(t_ccl[i], ms);
        }
    bi = -1;
    for (int i = 0; i < nc; i++)
        if (cand[i].thr_rx == pick.thr_rx && cand[i].thr_ry == pick.thr_ry &&
            (bi < 0 || t_ccl[i] < t_ccl[bi]))
            bi = i;
    if (bi >= 0) pick.ccl_rows = cand[bi].ccl_rows;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == ATQ_OK) *best = pick;
    return rc;
}

}  // namespace

extern "C" {

int atq_autotune(atq_detector *det, void *stream, const uint8_t *gray_in, size_t width_in,
                 size_t height_in, atq_shapes *best) {
    if (!det || !gray_in) return ATQ_ERR_BAD_ARG;
    const auto key = std::make_tuple(width_in, height_in, det->params.decimate);
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
        det->shapes = it->second;
    } else {
        FrontShapes sh;
        TRY(tune_frame(det, S(stream), gray_in, width_in, height_in, &sh));
        g_tuned[key] = sh;
        det->shapes = sh;
    }
    if (best) *best = atq_shapes{det->shapes.thr_rx, det->shapes.thr_ry, det->shapes.ccl_rows};
    return ATQ_OK;
}

int atq_detector_buffer(atq_detector *d, int which, void **ptr, size_t *capacity) {
    if (!d || !ptr || !capacity) return ATQ_ERR_BAD_ARG;
    const Caps &c = d->caps;
    const bool alt = d->sorted_in_alt;
    switch (which) {
    case ATQ_BUF_GRAY: *ptr = d->gray; *capacity = c.npx; break;
    case ATQ_BUF_THRESHOLDED: *ptr = d->thr; *capacity = c.npx; break;
    case ATQ_BUF_LABELS: *ptr = d->labels; *capacity = c.npx; break;
    case ATQ_BUF_SIZES: *ptr = d->sizes; *capacity = c.npx + 1; break;
    case ATQ_BUF_BOUNDARY_POINTS: *ptr = alt ? (void *)d->ws.vals_alt : (void *)d->vals; *capacity = c.cap; break;
    case ATQ_BUF_BOUNDARY_KEYS: *ptr = alt ? (void *)d->ws.keys_alt : (void *)d->keys; *capacity = c.cap; break;
    case ATQ_BUF_CLUSTER_BOUNDS: *ptr = d->bounds; *capacity = c.cap_cl; break;
    case ATQ_BUF_CLUSTER_POINTS: *ptr = d->cpts; *capacity = c.cap; break;
    case ATQ_BUF_CLUSTER_INDEX: *ptr = d->cidx; *capacity = c.cap; break;
    case ATQ_BUF_CLUSTER_EXTENTS: *ptr = d->cext; *capacity = c.cap_cl; break;
    case ATQ_BUF_LINE_FIT: *ptr = d->lf; *capacity = c.cap; break;
    case ATQ_BUF_CORNERS: *ptr = d->corners; *capacity = c.cap; break;
    case ATQ_BUF_PEAK_EXTENTS: *ptr = d->peaks; *capacity = c.cap_cl; break;
    case ATQ_BUF_FITTED: *ptr = d->fitted; *capacity = c.cap_cl; break;
    case ATQ_BUF_QUADS: *ptr = d->quads; *capacity = c.cap_cl; break;
    case ATQ_BUF_COUNTS: *ptr = d->counts; *capacity = 1; break;
    case ATQ_BUF_EXTREMA: *ptr = d->ext; *capacity = c.npx / 16; break;
    case ATQ_BUF_FILTERED_EXTREMA: *ptr = d->fil; *capacity = c.npx / 16; break;
    case ATQ_BUF_SPARSE_CORNERS: *ptr = d->sparse; *capacity = c.cap; break;
    case ATQ_BUF_PAIR_KEYS: *ptr = d->pair_keys; *capacity = c.pair_slots; break;
    case ATQ_BUF_PAIR_COUNTS: *ptr = d->pair_cnt; *capacity = c.pair_slots; break;
    default: return ATQ_ERR_BAD_ARG;
    }
    return ATQ_OK;
}

int atq_detector_set_timing(atq_detector *det, int enable) {
    if (!det) return ATQ_ERR_BAD_ARG;
    if (enable && !det->timing) {
        for (int i = 0; i <= ATQ_N_STAGES; i++)
            if (hipEventCreate(&det->ev[i]) != hipSuccess) return ATQ_ERR_HIP;
        det->timing = true;
    } else if (!enable && det->timing) {
        for (int i = 0; i <= ATQ_N_STAGES; i++) (void)hipEventDestroy(det->ev[i]);
        det->timing = false;
    }
    return ATQ_OK;
}

int atq_detector_stage_times(atq_detector *det, float *ms, int n) {
    if (!det || !ms || n < ATQ_N_STAGES || !det->timing) return ATQ_ERR_BAD_ARG;
    ATQ_HIP_CHECK(hipEventSynchronize(det->ev[ATQ_N_STAGES]));
    for (int i = 0; i < ATQ_N_STAGES; i++)
        ATQ_HIP_CHECK(hipEventElapsedTime(&ms[i], det->ev[i], det->ev[i + 1]));
    return ATQ_OK;
}

const char *atq_stage_name(int stage) {
    static const char *names[ATQ_N_STAGES] = {"threshold",  "ccl",       "boundary", "key_sort",
                                              "clusters",   "angle_fit", "fit_lines", "corners",
                                              "indexing",   "quads"};
    return (stage >= 0 && stage < ATQ_N_STAGES) ? names[stage] : "?";
}

int atq_detector_dims(atq_detector *det, size_t *width, size_t *height) {
    if (!det || !width || !height) return ATQ_ERR_BAD_ARG;
    *width = det->w;
    *height = det->h;
    return ATQ_OK;
}

}  // extern "C"

// ---- batched form ---------------------------------------------------------

struct atq_batch {
    std::vector<atq_detector *> dets;
    std::vector<uint8_t *> inputs;
    std::vector<size_t> w_in, h_in;
    static constexpr int N_STREAMS = 4;
    hipStream_t streams[N_STREAMS] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t joins[N_STREAMS] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool tuned = false;
    std::vector<int> stream_of;  // internal stream of every frame
};

namespace {

// Frames go to the internal streams longest first, each to the stream with
// the least work so far (work of a frame: its decimated pixels plus 0.5 M
// for the per-frame launches), which spreads the large frames over the
// streams and fills the gaps of their latency-bound kernels with the small
// ones (measured on the mixed 16-frame batch: slightly ahead of keeping the
// frames above 8 Mpx on one stream; 4 streams ahead of 2, 6 or 8).
void assign_streams(atq_batch *b, const size_t *w_in, const size_t *h_in,
                    const int32_t *decimate) {
    const size_t n = b->dets.size();
    const int ns = (int)std::min<size_t>(atq_batch::N_STREAMS, n);
    std::vector<double> cost(n), load(ns, 0.0);
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) {
        const int f = decimate ? std::max(decimate[i], 1) : 1;
        cost[i] = (double)(decimated(w_in[i], f) * decimated(h_in[i], f)) + 0.5e6;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t c) { return cost[a] > cost[c]; });
    b->stream_of.assign(n, 0);
    for (size_t i : order) {
        int best = 0;
        for (int k = 1; k < ns; k++)
            if (load[k] < load[best]) best = k;
        b->stream_of[i] = best;
        load[best] += cost[i];
    }
}

// On each internal stream, the thresholds and CCLs of its frames run as one
// launch chain per front-end shape (a table of frames per launch), then each
// frame's remaining stages; the streams are forked from and joined to `s`.
int batch_branches(atq_batch *b, hipStream_t s) {
    const size_t n = b->dets.size();
    const int ns = (int)std::min<size_t>(atq_batch::N_STREAMS, n);
    for (size_t i = 0; i < n; i++) TRY(frame_setup(b->dets[i], b->w_in[i], b->h_in[i]));
    ATQ_HIP_CHECK(hipEventRecord(b->fork, s));
    for (int k = 0; k < ns; k++) {
        hipStream_t sk = b->streams[k];
        ATQ_HIP_CHECK(hipStreamWaitEvent(sk, b->fork, 0));
        std::vector<bool> done(n, false);
        for (size_t i = 0; i < n; i++) {
            if (done[i] || b->stream_of[i] != k) continue;
            const FrontShapes sh = b->dets[i]->shapes;
            const int mode = b->dets[i]->params.mode;
            FrontFrame group[FRONT_MAX_FRAMES];
            int g = 0;
            for (size_t j = i; j < n && g < FRONT_MAX_FRAMES; j++) {
                const FrontShapes &o = b->dets[j]->shapes;
                if (done[j] || b->stream_of[j] != k || o.thr_rx != sh.thr_rx ||
                    o.thr_ry != sh.thr_ry || o.ccl_rows != sh.ccl_rows ||
                    b->dets[j]->params.mode != mode)
                    continue;
                group[g++] = front_frame(b->dets[j], b->inputs[j], b->w_in[j], b->h_in[j]);
                done[j] = true;
            }
            TRY(launch_threshold_frames(sk, group, g, sh));
            if (mode == ATQ_MODE_FIXED) TRY(launch_ccl_fixed_frames(sk, group, g));
            else TRY(launch_ccl_frames(sk, group, g, sh));
        }
        for (size_t i = 0; i < n; i++)
            if (b->stream_of[i] == k) TRY(run_back(b->dets[i], sk, b->inputs[i]));
        ATQ_HIP_CHECK(hipEventRecord(b->joins[k], sk));
        ATQ_HIP_CHECK(hipStreamWaitEvent(s, b->joins[k], 0));
    }
    return ATQ_OK;
}

}  // namespace

extern "C" {

int atq_batch_destroy(atq_batch *b) {
    if (!b) return ATQ_ERR_BAD_ARG;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto *d : b->dets)
        if (d) atq_detector_destroy(d);
    for (auto *p : b->inputs)
        if (p) (void)hipFree(p);
    for (int k = 0; k < atq_batch::N_STREAMS; k++) {
        if (b->streams[k]) (void)hipStreamDestroy(b->streams[k]);
        if (b->joins[k]) (void)hipEventDestroy(b->joins[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
    return ATQ_OK;
}

int atq_batch_create(size_t n_frames, const size_t *widths_in, const size_t *heights_in,
                     const int32_t *decimate, atq_batch **out) {
    if (!out || !widths_in || !heights_in || n_frames == 0) return ATQ_ERR_BAD_ARG;
    *out = nullptr;
    if (int r = atq_init()) return r;
    atq_batch *b = new (std::nothrow) atq_batch();
    if (!b) return ATQ_ERR_BAD_ARG;
    int rc = ATQ_OK;
    for (size_t i = 0; i < n_frames && rc == ATQ_OK; i++) {
        atq_params p{decimate ? decimate[i] : 1, 5, 25, 8, 0};
        atq_detector *d = nullptr;
        rc = atq_detector_create(widths_in[i], heights_in[i], &p, &d);
        if (d) d->in_batch = true;
        b->dets.push_back(d);
        uint8_t *in = nullptr;
        if (rc == ATQ_OK && hipMalloc(&in, widths_in[i] * heights_in[i]) != hipSuccess) rc = ATQ_ERR_HIP;
        b->inputs.push_back(in);
        b->w_in.push_back(widths_in[i]);
        b->h_in.push_back(heights_in[i]);
    }
    for (int k = 0; k < atq_batch::N_STREAMS && rc == ATQ_OK; k++) {
        if (hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->joins[k], hipEventDisableTiming) != hipSuccess)
            rc = ATQ_ERR_HIP;
    }
    if (rc == ATQ_OK && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        rc = ATQ_ERR_HIP;
    if (rc == ATQ_OK) assign_streams(b, widths_in, heights_in, decimate);
    if (rc != ATQ_OK) {
        atq_batch_destroy(b);
        return rc;
    }
    *out = b;
    return ATQ_OK;
}

int atq_batch_input(atq_batch *b, size_t i, uint8_t **gray_in) {
    if (!b || !gray_in || i >= b->inputs.size()) return ATQ_ERR_BAD_ARG;
    *gray_in = b->inputs[i];
    return ATQ_OK;
}

int atq_batch_frame(atq_batch *b, size_t i, atq_detector **det) {
    if (!b || !det || i >= b->dets.size()) return ATQ_ERR_BAD_ARG;
    *det = b->dets[i];
    return ATQ_OK;
}

int atq_batch_autotune(atq_batch *b, void *stream) {
    if (!b) return ATQ_ERR_BAD_ARG;
    for (size_t i = 0; i < b->dets.size(); i++)
        TRY(atq_autotune(b->dets[i], stream, b->inputs[i], b->w_in[i], b->h_in[i], nullptr));
    b->tuned = true;
    return ATQ_OK;
}

int atq_batch_detect(atq_batch *b, void *stream, int use_graph) {
    if (!b) return ATQ_ERR_BAD_ARG;
    hipStream_t s = S(stream);
    if (!b->tuned) TRY(atq_batch_autotune(b, stream));
    if (!use_graph) return batch_branches(b, s);
    if (!s) return ATQ_ERR_BAD_ARG;
    if (!b->exec) {
        for (auto *d : b->dets)
            if (d->timing) return ATQ_ERR_BAD_ARG;  // no timing events inside a graph
        ATQ_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = batch_branches(b, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        if (rc != ATQ_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return ATQ_ERR_HIP;
        b->graph = g;
        ATQ_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    ATQ_HIP_CHECK(hipGraphLaunch(b->exec, s));
    return ATQ_OK;
}

}  // extern "C"
