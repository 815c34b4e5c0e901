/*
 * atq.h - C ABI of the MI355X AprilTag quad front-end (threshold -> quads).
 *
 * Every entry point is `extern "C"`, takes plain device pointers and sizes,
 * returns an int status (ATQ_OK = 0) and never throws. `stream` is a
 * hipStream_t passed as `void*` (NULL = the default stream); all work is
 * ordered on it. Entry points that mirror a reference stage name the
 * reference declaration they replace (paths relative to the reference repo
 * carbotaniuman/sycl-apriltag-test).
 *
 * Record types are layout-identical to the reference's (static_asserts in
 * the library). The one widening: cluster indices are uint32_t instead of
 * uint16_t (the reference's u16 index and its 65535 sentinel cap the frame at
 * 65535 clusters, src/main.cpp:639-651, include/fit_quad.h:23). `atq_corner`
 * keeps its 12-byte size because the reference struct already pads the u16
 * to 4 bytes.
 */
#ifndef ATQ_H
#define ATQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    ATQ_OK = 0,
    ATQ_ERR_BAD_DIMS = 1,  /* width/height not multiples of 4, < 8 or > 8192 */
    ATQ_ERR_CAPACITY = 2,  /* a count exceeded the buffer it must fit in     */
    ATQ_ERR_HIP = 3,       /* a HIP runtime call failed                      */
    ATQ_ERR_BAD_ARG = 4,   /* NULL pointer, bad factor, too-small workspace  */
    ATQ_ERR_NO_DEVICE = 5  /* no gfx950 device visible                       */
};

/* ---- records (layout-identical to the reference) ------------------------ */

/* sycl::vec<uint8_t,2> of tile {min,max}: include/threshold.h:7-8 */
typedef struct atq_extrema {
    uint8_t min, max;
} atq_extrema;

/* include/boundary_point.h:42-51. packed_x = [half_pixel:2 | x:14],
 * packed_y = [black_to_white:1 | unused:1 | y:14] */
typedef struct atq_boundary_point {
    uint16_t packed_x, packed_y;
} atq_boundary_point;

/* include/boundary_point.h:100-105 */
typedef struct atq_cluster_point {
    uint16_t packed_x, packed_y;
    float slope;
} atq_cluster_point;

/* include/cluster_bounds.h:9-12 */
typedef struct atq_cluster_extents {
    uint32_t start, count;
} atq_cluster_extents;

/* include/cluster_bounds.h:14-26 */
typedef struct atq_cluster_bounds {
    uint16_t x_min, x_max, y_min, y_max;
    int32_t gx_sum, gy_sum;
    int64_t pxgx_pygy_sum;
    uint32_t start, count;
} atq_cluster_bounds;

/* include/line_fit_point.h:11-17 (Mx,My doubled; Mxx,Myy,Mxy quadrupled) */
typedef struct atq_line_fit_point {
    int64_t Mx, My, Mxx, Myy, Mxy, W, a, b;
} atq_line_fit_point;

/* include/fit_quad.h:9-12 */
typedef struct atq_peak_extents {
    uint32_t start, count;
} atq_peak_extents;

/* include/fit_quad.h:21-25, cluster_index widened u16 -> u32 (same size) */
typedef struct atq_corner {
    uint32_t line_fit_point_index;
    uint32_t cluster_index;
    float error;
} atq_corner;

/* include/fit_quad.h:27-31 */
typedef struct atq_fitted_quad {
    atq_line_fit_point moments[4];
    uint16_t num_in_moments[4];
    uint16_t indices[4];
} atq_fitted_quad;

/* include/fit_quad.h:33-37. blob_index (left unset by the reference,
 * src/main.cpp:1049) carries the peak-segment index the quad came from. */
typedef struct atq_quad_corners {
    float corners[4][2];
    uint8_t reversed_border;
    uint8_t pad_[3];
    uint32_t blob_index;
} atq_quad_corners;

/* Compile-time constants of the reference, in one struct (defaults equal
 * the reference; only the compat values are supported for parity). */
typedef struct atq_params {
    int32_t decimate;            /* point subsample factor; 1 = none        */
    int32_t min_white_black_diff;/* src/threshold.cpp:86 (5)                */
    int32_t min_component_size;  /* src/find_boundaries.cpp:82 (25)         */
    int32_t min_tag_width;       /* src/main.cpp:1037 (8)                   */
    int32_t pair_table_log2;     /* detector only: its cluster-pair table has
                                    2^n slots, n in [8, 24]; 0 = automatic
                                    (about one slot per 64 pixels, 2^14 to
                                    2^20). No counterpart in the reference. */
    int32_t mode;                /* detector only: ATQ_MODE_COMPAT (0, the
                                    reference's results) or ATQ_MODE_FIXED */
} atq_params;

/* Detector modes. COMPAT reproduces the reference bit for bit. FIXED corrects
 * seven of its behaviours (no reference counterpart; never used for parity):
 *  - CCL is pixel-level, white 8-connected, black 4-connected, with no links
 *    through the reference's colour-agnostic neighbour mask
 *    (src/segmentation.cpp:139-255). Labels are root + 1 (white with the
 *    0x80000000 bit), 0 for 127 pixels; ATQ_BUF_SIZES holds the count of the
 *    component with root r at index r + 1 and 0 at index 0, so neither 127
 *    pixels nor out-of-image neighbours form boundary records
 *    (src/main.cpp:310-311, src/find_boundaries.cpp:49-55);
 *  - column 0 is a local pixel of the boundary pass (src/find_boundaries.cpp:66-69);
 *  - records decode both half-pixel bits: doubled coordinates 2x + dx,
 *    2y + dy and gradient +-(dx, dy) of the plane's neighbour offset
 *    (include/boundary_point.h:16);
 *  - the line-fit weight samples pixel (rx/2, ry/2) (include/line_fit_point.h:63);
 *  - fit_lines considers global point 0 like every other (src/fit_quad.cpp:184-187);
 *  - do_indexing fits each peak segment to its own cluster (src/fit_quad.cpp:405-408). */
enum { ATQ_MODE_COMPAT = 0, ATQ_MODE_FIXED = 1 };

/* Per-frame counts written by the pipeline (device memory, 16 x u32). */
typedef struct atq_counts {
    uint32_t n_boundary;      /* N_bp: boundary records                    */
    uint32_t n_pairs;         /* cluster-pair segments                     */
    uint32_t n_clusters;      /* clusters passing ValidBlobFilter          */
    uint32_t n_cluster_points;/* points in those clusters                  */
    uint32_t n_corners;       /* corners with error != 0                   */
    uint32_t n_segments;      /* peak segments E                           */
    uint32_t n_quads;         /* quads surviving process_fitted_quads      */
    uint32_t overflow;        /* bit flags; non-zero = the frame FAILED and
                                 its results are void: 1 boundary buffer,
                                 2 corner buffer, 4 a bounded wait (a
                                 look-back scan or a grid barrier) gave up.
                                 Callers must check it before using any
                                 result of the frame                       */
    uint32_t oob_reads;       /* do_indexing reads past the moment array   */
    uint32_t key_label_bits;  /* lb of the detector's packed pair keys     */
    uint32_t pair_filter_off; /* 1: the detector kept every boundary record
                                 (its pair table overflowed); 0: it dropped
                                 the records of pairs whose record count
                                 ValidBlobFilter rejects (< 24 or
                                 > 2 * (2w + 2h))                          */
    uint32_t n_valid_pairs;   /* detector: cluster pairs passing
                                 ValidBlobFilter (from its pair table)     */
    uint32_t sort_passes;     /* detector: radix passes over the pair-rank
                                 keys (1 or 2); 0 = the full-key sort ran
                                 (pair table overflow or too many valid
                                 pairs to rank)                            */
    uint32_t rank_shift;      /* detector: 2 = rank keys carry the plane in
                                 2 low bits (records in raster order); 0 =
                                 records reach the sort plane by plane     */
    uint32_t reserved[2];
} atq_counts;

/* ---- library ------------------------------------------------------------ */

/* Checks that device 0 (or the current device) is gfx950 and records it. */
int atq_init(void);
/* Human-readable text of a status code. */
const char *atq_status_string(int status);

/* ---- stage drop-ins ----------------------------------------------------- */

/* Point-subsample decimation out[y][x] = in[f*y][f*x],
 * out dims 1+(w-1)/f x 1+(h-1)/f. No reference counterpart (notes.txt:18-21
 * says decimation is "Ignored for now"); follows upstream apriltag's
 * image_u8_decimate for integer factors. */
int atq_decimate(void *stream, const uint8_t *gray_in, size_t width_in,
                 size_t height_in, int factor, uint8_t *gray_out);

/* Colour removal plus decimation (notes.txt:19-22, "1. Remove Colors
 * 2. Decimate"): out[y][x] = grey of in[f*y][f*x] for interleaved 8-bit
 * pixels of `channels` bytes (1 grey, 3 RGB, 4 RGBA). The grey of a colour
 * pixel is the one the reference's loader computes (stbi_load with
 * STBI_grey, src/main.cpp:265: stbi__compute_y, include/stb_image.h:1834-1836,
 * (77 r + 150 g + 29 b) >> 8, alpha dropped). */
int atq_grey_decimate(void *stream, const uint8_t *pixels_in, size_t width_in,
                      size_t height_in, int channels, int factor,
                      uint8_t *gray_out);

/* Replaces threshold_image(q, grayscale, extrema, filtered_extrema,
 * thresholded, width, height, deps)  include/threshold.h:6-10.
 * Writes all three outputs like the reference (src/threshold.cpp:8-95). */
int atq_threshold_image(void *stream, const uint8_t *grayscale,
                        atq_extrema *extrema, atq_extrema *filtered_extrema,
                        uint8_t *thresholded, size_t width, size_t height);

/* Fused decimation + threshold: decimate by `factor` (1 = none) and threshold in one
 * pass. gray_out receives the decimated image when factor > 1 (it feeds the
 * line-fit weights) and may be NULL when factor == 1. extrema and
 * filtered_extrema may be NULL. */
int atq_decimate_threshold(void *stream, const uint8_t *gray_in,
                           size_t width_in, size_t height_in, int factor,
                           uint8_t *gray_out, atq_extrema *extrema,
                           atq_extrema *filtered_extrema,
                           uint8_t *thresholded);

/* Replaces image_segmentation(q, thresholded, label_scratch, labels, sizes,
 * sizes_elem, width, height, deps)  include/segmentation.h:11-15.
 * labels: white = 0x80000000|root, black = root, 127 -> 0, root = the
 * component's minimum node id (src/segmentation.cpp:348-388). sizes (width*height entries) is
 * written at every label root and at index 0; other entries are not read by
 * any stage and are left as they were. sizes_elem is accepted for signature
 * parity and, as in the reference, unused. label_scratch (width*height u32)
 * is workspace. */
int atq_image_segmentation(void *stream, const uint8_t *thresholded,
                           uint32_t *label_scratch, uint32_t *labels,
                           uint32_t *sizes, size_t sizes_elem, size_t width,
                           size_t height);

/* Largest frame (width * height decimated pixels) that
 * atq_image_segmentation_static serves. */
#define ATQ_SEGMENTATION_STATIC_MAX_PIXELS 4194304u

/* atq_image_segmentation without label_scratch, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The scratch is a 12 MiB array in the library's device memory, one per GPU
 * the library runs on, held until the process ends; this call allocates
 * nothing. Results equal atq_image_segmentation's. Frames of more than
 * ATQ_SEGMENTATION_STATIC_MAX_PIXELS pixels return ATQ_ERR_CAPACITY. Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams or host threads) race and their results are void;
 * order them on one stream. */
int atq_image_segmentation_static(void *stream, const uint8_t *thresholded,
                                  uint32_t *labels, uint32_t *sizes,
                                  size_t sizes_elem, size_t width,
                                  size_t height);

/* Replaces find_boundaries(q, labels, sizes, points, blob_labels, width,
 * height, deps)  include/find_boundaries.h:8-12. Writes the 4 sparse planes
 * of 4*width*height entries exactly as the reference leaves them after its
 * pre-fill (points 0, keys all-ones where nothing is emitted,
 * src/main.cpp:312-315); the pre-fill is done here. */
int atq_find_boundaries(void *stream, const uint32_t *labels,
                        const uint32_t *sizes, atq_boundary_point *points,
                        uint64_t *blob_labels, size_t width, size_t height);

/* ---- glue between stages (inline oneDPL calls in src/main.cpp) ---------- */

/* Bytes of workspace the glue entry points below need for a frame of at
 * most width x height (decimated) pixels. */
size_t atq_glue_workspace_size(size_t width, size_t height);

/* Compaction + key sort (src/main.cpp:467-477, 515-518): compact the sparse planes (drop
 * points whose u32 image is 0) and group by cluster-pair key ascending;
 * within a key the order is plane-major raster, i.e. a stable sort.
 * *count_dev (device u32) receives N_bp. Outputs hold 4*width*height. */
int atq_group_boundaries(void *stream, const atq_boundary_point *points,
                         const uint64_t *blob_labels, size_t width,
                         size_t height, atq_boundary_point *out_points,
                         uint64_t *out_keys, uint32_t *count_dev,
                         void *workspace, size_t workspace_bytes);

/* Cluster bounds + filter (src/main.cpp:571-609): per key segment ClusterBounds
 * (include/cluster_bounds.h:28-83) then ValidBlobFilter(width, height)
 * (include/cluster_bounds.h:85-120). n_points is read from device memory.
 * out_bounds receives the filtered bounds, counts->n_pairs/n_clusters set. */
int atq_cluster_bounds_filter(void *stream, const atq_boundary_point *points,
                              const uint64_t *keys, const uint32_t *n_points_dev,
                              size_t width, size_t height,
                              atq_cluster_bounds *out_bounds,
                              atq_counts *counts_dev, void *workspace,
                              size_t workspace_bytes);

/* Cluster assignment, extents, angle sort (src/main.cpp:621-750): points of valid clusters with their
 * calc_theta slope and cluster index, cluster extents, sorted by (cluster,
 * slope) stably. */
int atq_cluster_points_sort(void *stream, const atq_boundary_point *points,
                            const uint32_t *n_points_dev,
                            const atq_cluster_bounds *bounds,
                            atq_counts *counts_dev,
                            atq_cluster_point *out_points,
                            uint32_t *out_cluster_index,
                            atq_cluster_extents *out_extents, void *workspace,
                            size_t workspace_bytes);

/* Line-fit points + moment prefix (src/main.cpp:768-801): compute_initial_linefit
 * (include/line_fit_point.h:50-86) and the per-cluster inclusive prefix sum
 * of the moments (exact int64). The clusters are the `extents` runs of the
 * angle-sorted points (the reference keys the segmented scan by the
 * per-point cluster index, whose runs are these extents);
 * counts_dev->n_clusters gives their number. */
int atq_linefit_prefix(void *stream, const atq_cluster_point *points,
                       const atq_cluster_extents *extents,
                       const atq_counts *counts_dev, const uint8_t *grayscale,
                       size_t width, size_t height, atq_line_fit_point *out);

/* Replaces fit_lines(q, points, cluster_indices, cluster_extents,
 * points_size, found_corners, deps)  include/fit_quad.h:46-50.
 * Blocking like the reference (src/fit_quad.cpp:374). found_corners
 * (points_size entries) is zero-filled here, then corners are written. */
int atq_fit_lines(void *stream, const atq_line_fit_point *points,
                  const uint32_t *cluster_indices,
                  const atq_cluster_extents *cluster_extents,
                  size_t points_size, atq_corner *found_corners);

/* Corner compaction, sort, segments (src/main.cpp:832-907): compact corners with error != 0, sort by
 * (cluster asc, error desc) stably, segment per cluster. */
int atq_corner_segments(void *stream, const atq_corner *found_corners,
                        atq_counts *counts_dev, atq_corner *compacted,
                        atq_peak_extents *segments, void *workspace,
                        size_t workspace_bytes);

/* Replaces do_indexing(q, extents, extents_count, compacted_corners,
 * points, cluster_extents, fitted)  include/fit_quad.h:52-55.
 * Blocking like the reference (src/fit_quad.cpp:534). fitted
 * (extents_count entries) is zero-filled here first. n_points bounds the
 * moment array (reads past it, possible because segment e is paired with
 * cluster_extents[e] as in the reference, read zeros). */
int atq_do_indexing(void *stream, const atq_peak_extents *extents,
                    size_t extents_count, const atq_corner *compacted_corners,
                    const atq_line_fit_point *points, size_t n_points,
                    const atq_cluster_extents *cluster_extents,
                    atq_fitted_quad *fitted);

/* process_fitted_quads (src/main.cpp:1035-1162) minus tag decoding:
 * refit, intersect, area/angle/convexity checks; survivors are written in
 * input order. *n_quads_dev (device u32) receives the count. */
int atq_quads_from_fits(void *stream, const atq_fitted_quad *fitted,
                        size_t n_fitted, atq_quad_corners *out,
                        uint32_t *n_quads_dev, void *workspace,
                        size_t workspace_bytes);

/* ---- differentiable line-fit / corner stage ----------------------------- */
/* No reference counterpart. The reference's line-fit weight is an integer
 * read at the wrong row (include/line_fit_point.h:57); this form weights
 * point i by w_i = s_i * sqrt(1 + gx^2 + gy^2), (gx, gy) the central
 * differences of a float image at the point's pixel (x+1, y), s_i a per-point
 * boundary weight, so the corners are differentiable in the pixels and in s.
 * A quad is four lines fitted (weighted centroid + major axis, float32 with
 * compensated sums) to the cyclic point ranges [idx0,idx1], [idx1,idx2],
 * [idx2,idx3], [idx3,idx0] of one angle-sorted cluster - the ranges
 * do_indexing forms - and corner k intersects lines k and k+1 mod 4, as
 * process_fitted_quads does (src/main.cpp:1059-1096). */
typedef struct atq_diff_quad {
    uint32_t cluster_start, cluster_count; /* the cluster's extents in points */
    uint32_t idx[4];                       /* ascending indices in the cluster */
} atq_diff_quad;

/* corners: n_quads x 4 x 2 floats. boundary_weights (n_points) may be NULL
 * (all 1). w_out (n_points) may be NULL; else receives w_i of every point of
 * every quad's cluster. */
int atq_linefit_corners_forward(void *stream, const float *image, size_t width,
                                size_t height, const atq_cluster_point *points,
                                size_t n_points, const float *boundary_weights,
                                const atq_diff_quad *quads, size_t n_quads,
                                float *corners, float *w_out);

/* Vector-Jacobian product of the forward: given grad_corners = dL/dcorners
 * (n_quads x 8), ADDS dL/d(boundary weight) into grad_weights (n_points) and
 * dL/d(pixel) into grad_image (width*height); either may be NULL. Callers
 * zero them first. */
int atq_linefit_corners_backward(void *stream, const float *image,
                                 size_t width, size_t height,
                                 const atq_cluster_point *points,
                                 size_t n_points,
                                 const float *boundary_weights,
                                 const atq_diff_quad *quads, size_t n_quads,
                                 const float *grad_corners,
                                 float *grad_weights, float *grad_image);

/* ---- tag payload decode ------------------------------------------------- */
/* The table-free part of the decode the reference runs on every quad:
 * decode_quads (src/quad_decoding.cpp:77-128) hands each one to the apriltag
 * library's quad_decode_task (src/quad_decoding.cpp:46-75), a build-time
 * dependency (carbotaniuman/apriltag @ a8b2f96, CMakeLists.txt:11-17) absent
 * from the reference checkout together with its tag36h11 family. What needs
 * no family table is built here, after apriltag's quad_decode: the
 * homography from the tag square [-1,1]^2 (corner i at (-1,-1), (1,-1),
 * (1,1), (-1,1)) to the quad; least-squares planes of the gray level sampled
 * just outside (white) and just inside (black) the border; the polarity
 * test; one bilinear sample per bit cell against the mean of the two planes,
 * sharpened by a 3x3 Laplacian; the code word and its decision margin. The
 * family's bit layout and code table come from the caller. The image is the
 * one the quads were found in (the decimated grey image at decimate > 1). */
#define ATQ_TAG_MAX_BITS 64
#define ATQ_TAG_MAX_WIDTH 16
typedef struct atq_tag_layout {
    int32_t width_at_border;  /* cells across the outer edge of the black
                                 border, 2 .. ATQ_TAG_MAX_WIDTH            */
    int32_t total_width;      /* cells across the whole tag, width_at_border
                                 .. ATQ_TAG_MAX_WIDTH, same parity         */
    int32_t reversed_border;  /* 1: a white border inside a black ring     */
    int32_t nbits;            /* 1 .. ATQ_TAG_MAX_BITS                     */
    float decode_sharpening;  /* apriltag's default is 0.25                */
    uint8_t bit_x[ATQ_TAG_MAX_BITS]; /* cell of bit i (bit 0 the most      */
    uint8_t bit_y[ATQ_TAG_MAX_BITS]; /* significant), border coordinates:  */
                                     /* 0 .. width_at_border - 1 is the    */
                                     /* black border square                */
} atq_tag_layout;

typedef struct atq_tag_code {
    uint64_t code;           /* nbits read in the quad's corner order        */
    float decision_margin;   /* min(mean white, mean black bit margin);
                                -1 (and code 0): the polarity test failed    */
    float homography[9];     /* tag square -> image, row-major, [8] = 1      */
} atq_tag_code;

/* Decodes quads[0 .. n) (device) in `gray` (device, width x height u8) to
 * out[0 .. n) (device). n = n_quads, or min(*n_quads_dev, n_quads) when
 * n_quads_dev (device u32, e.g. the detector's counts) is non-null. layout is
 * a host pointer, read before the call returns. Stream-ordered. */
int atq_decode_payloads(void *stream, const uint8_t *gray, size_t width,
                        size_t height, const atq_quad_corners *quads,
                        size_t n_quads, const uint32_t *n_quads_dev,
                        const atq_tag_layout *layout, atq_tag_code *out);

typedef struct atq_tag_match {
    int32_t id;        /* index in the table; -1: none within max_hamming  */
    int32_t hamming;   /* bits that differ                                 */
    int32_t rotation;  /* quarter turns of the read code, each taking the
                          cell (x, y) to (width_at_border - 1 - y, x)      */
    int32_t pad_;
} atq_tag_match;

/* The nearest code of table[0 .. n_table) (device u64, the family's codes in
 * the layout's bit order) over the four rotations of each codes[i].code,
 * ties to the lower rotation, then the lower index (the role of apriltag's
 * quick_decode, which looks the code up in the family's hash of codes within
 * its Hamming limit). Codes whose polarity test failed match nothing. n and
 * n_codes_dev as for atq_decode_payloads; the layout's cells must be closed
 * under the quarter turn (ATQ_ERR_BAD_ARG otherwise). */
int atq_match_codes(void *stream, const atq_tag_code *codes, size_t n_codes,
                    const uint32_t *n_codes_dev, const uint64_t *table,
                    size_t n_table, const atq_tag_layout *layout,
                    int max_hamming, atq_tag_match *out);

/* ---- whole pipeline ----------------------------------------------------- */

typedef struct atq_detector atq_detector;

/* Allocates every buffer for frames up to max_width_in x max_height_in
 * (before decimation) once; atq_detect_quads then never allocates. */
int atq_detector_create(size_t max_width_in, size_t max_height_in,
                        const atq_params *params, atq_detector **out);
int atq_detector_destroy(atq_detector *det);

/* threshold -> quads for one frame already in device memory. Fully
 * stream-ordered: no host synchronisation. Quads land in the detector
 * (atq_detector_buffer ATQ_BUF_QUADS) and counts in ATQ_BUF_COUNTS. */
int atq_detect_quads(atq_detector *det, void *stream, const uint8_t *gray_in,
                     size_t width_in, size_t height_in);
/* The same from interleaved RGB (channels 3) or RGBA (4) pixels: the grey
 * conversion of atq_grey_decimate is fused into the threshold kernel's loads,
 * and the (decimated) grey image lands in ATQ_BUF_GRAY. Results equal
 * atq_detect_quads on atq_grey_decimate's output. No reference counterpart
 * (the reference converts on the host while loading, src/main.cpp:265). */
int atq_detect_quads_rgb(atq_detector *det, void *stream,
                         const uint8_t *pixels_in, size_t width_in,
                         size_t height_in, int channels);

enum {
    ATQ_BUF_GRAY = 0,        /* u8  decimated grey; not written for an
                                undecimated grey frame (read in place) */
    ATQ_BUF_THRESHOLDED = 1, /* u8  W*H                                  */
    ATQ_BUF_LABELS = 2,      /* u32 W*H                                  */
    ATQ_BUF_SIZES = 3,       /* u32 W*H + 1; at a root its component's
                                pixel count (fixed mode: at root + 1), exact
                                below min_component_size and only
                                ">= min_component_size" above            */
    ATQ_BUF_BOUNDARY_POINTS = 4, /* atq_boundary_point, grouped by key   */
    ATQ_BUF_BOUNDARY_KEYS = 5,   /* u64 packed min<<lb|max, lb = counts.key_label_bits */
    ATQ_BUF_CLUSTER_BOUNDS = 6,  /* filtered atq_cluster_bounds; start
                                    indexes BOUNDARY_POINTS, which holds
                                    only the records of the pairs that pass
                                    ValidBlobFilter (see pair_filter_off) */
    ATQ_BUF_CLUSTER_POINTS = 7,  /* atq_cluster_point, angle-sorted      */
    ATQ_BUF_CLUSTER_INDEX = 8,   /* u32                                  */
    ATQ_BUF_CLUSTER_EXTENTS = 9, /* atq_cluster_extents                  */
    ATQ_BUF_LINE_FIT = 10,       /* atq_line_fit_point prefix sums       */
    ATQ_BUF_CORNERS = 11,        /* atq_corner compacted + sorted        */
    ATQ_BUF_PEAK_EXTENTS = 12,   /* atq_peak_extents                     */
    ATQ_BUF_FITTED = 13,         /* atq_fitted_quad                      */
    ATQ_BUF_QUADS = 14,          /* atq_quad_corners                     */
    ATQ_BUF_COUNTS = 15,         /* atq_counts                           */
    ATQ_BUF_EXTREMA = 16,        /* atq_extrema (W/4)*(H/4); not written
                                    by atq_detect_quads (no later stage
                                    reads the tile extrema: the drop-in
                                    atq_threshold_image writes them)     */
    ATQ_BUF_FILTERED_EXTREMA = 17, /* as ATQ_BUF_EXTREMA                 */
    ATQ_BUF_SPARSE_CORNERS = 18, /* atq_corner, one slot per cluster point */
    ATQ_BUF_PAIR_KEYS = 19,      /* u64 ~(min<<32|max) per pair-table slot, 0 = empty */
    ATQ_BUF_PAIR_COUNTS = 20     /* u32 record count per slot (exact up to
                                    2 * (2w + 2h), above it only "greater") */
};
/* Device pointer and capacity (in elements) of a detector buffer. */
int atq_detector_buffer(atq_detector *det, int which, void **ptr,
                        size_t *capacity);
/* Decimated dims of the last frame. */
int atq_detector_dims(atq_detector *det, size_t *width, size_t *height);

/* Per-stage timing, the counterpart of the reference's `prog` prints of
 * cumulative microseconds after each stage (src/main.cpp:330-335 ...
 * 972-977). When enabled, atq_detect_quads records a HIP event on its stream
 * after each stage; atq_detector_stage_times waits for the last frame's
 * events and returns each stage's milliseconds. */
enum {
    ATQ_STAGE_THRESHOLD = 0,   /* decimate + tile extrema + threshold       */
    ATQ_STAGE_CCL = 1,         /* union-find labelling + sizes              */
    ATQ_STAGE_BOUNDARY = 2,    /* boundary records                          */
    ATQ_STAGE_KEY_SORT = 3,    /* radix sort by cluster-pair key            */
    ATQ_STAGE_CLUSTERS = 4,    /* segments, bounds, filter, extents, slopes */
    ATQ_STAGE_ANGLE_FIT = 5,   /* angle sort + line-fit moments + prefix    */
    ATQ_STAGE_FIT_LINES = 6,   /* corner candidates                         */
    ATQ_STAGE_CORNERS = 7,     /* corner grouping + sort                    */
    ATQ_STAGE_INDEXING = 8,    /* 4-corner selection                        */
    ATQ_STAGE_QUADS = 9,       /* corner geometry + checks                  */
    ATQ_N_STAGES = 10
};
int atq_detector_set_timing(atq_detector *det, int enable);

/* ---- front-end shapes and the autotuner ----------------------------------- */
/* No reference counterpart (the reference fixes its work-group shapes).
 * Threshold regions of threshold_tiles_x x threshold_tiles_y 4x4-pixel tiles
 * (64x16, 32x32 or 16x64) and CCL tiles of 32 x ccl_block_rows 2x2-pixel
 * blocks (8, 16 or 32). Results do not depend on the shapes; the default is
 * 64x16 and 16. */
typedef struct atq_shapes {
    int32_t threshold_tiles_x, threshold_tiles_y, ccl_block_rows;
} atq_shapes;
int atq_detector_set_shapes(atq_detector *det, const atq_shapes *shapes);
int atq_detector_get_shapes(atq_detector *det, atq_shapes *shapes);
/* Times every supported shape on gray_in (a width_in x height_in frame in
 * device memory) with det's buffers (their contents are overwritten), keeps
 * the fastest threshold and CCL shapes in det and in the library's table for
 * that (width_in, height_in, decimate), and returns them in *best (may be
 * NULL). Blocks until the timing runs end. A resolution already in the table
 * takes its shapes from there without timing. */
int atq_autotune(atq_detector *det, void *stream, const uint8_t *gray_in,
                 size_t width_in, size_t height_in, atq_shapes *best);
int atq_detector_stage_times(atq_detector *det, float *ms, int n);
const char *atq_stage_name(int stage);

/* ---- batched form --------------------------------------------------------- */
/* No reference counterpart. Many frames of uneven sizes and decimation
 * factors through one launch chain: each frame has its own detector (sized
 * for it) and input buffer. The thresholds and CCLs of all frames run as
 * one persistent launch per kernel over a table of frames (one chain per
 * autotuned shape), the remaining stages of the frames as parallel branches
 * over internal streams; the whole chain is captured once into a HIP graph,
 * so a batch is one graph launch. Results per frame are bit-identical to
 * atq_detect_quads on that frame and are read through atq_batch_frame's
 * detector. */
typedef struct atq_batch atq_batch;
int atq_batch_create(size_t n_frames, const size_t *widths_in,
                     const size_t *heights_in, const int32_t *decimate,
                     atq_batch **out);
int atq_batch_destroy(atq_batch *batch);
/* Device input buffer (width_in*height_in bytes) of frame i. */
int atq_batch_input(atq_batch *batch, size_t i, uint8_t **gray_in);
/* Runs every frame. use_graph != 0: launches the captured graph (capturing
 * it on the first call); 0: the same branches launched eagerly. Ordered on
 * `stream` (which must not be the NULL stream when use_graph != 0). */
int atq_batch_detect(atq_batch *batch, void *stream, int use_graph);
/* Autotunes every frame's shapes on its current input (atq_autotune; each
 * resolution is timed once). atq_batch_detect does this on its first call
 * unless it was done before. */
int atq_batch_autotune(atq_batch *batch, void *stream);
int atq_batch_frame(atq_batch *batch, size_t i, atq_detector **det);

#ifdef __cplusplus
}
#endif
#endif /* ATQ_H */
