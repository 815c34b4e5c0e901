// api.cpp - the C ABI (include/atq.h): argument checks, stage drop-ins.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "kernels.hpp"
#include "pipeline.hpp"

using namespace atq;

namespace {

bool g_ready = false;

// Scratch of atq_image_segmentation_static: the CCL's parent and state words
// (3/4 word per pixel) for frames up to ATQ_SEGMENTATION_STATIC_MAX_PIXELS.
// Every word the CCL reads is written earlier in the same call, so nothing
// needs clearing between calls.
constexpr size_t STATIC_SCRATCH_WORDS = (size_t)ATQ_SEGMENTATION_STATIC_MAX_PIXELS / 4 * 3;
__device__ __align__(256) uint32_t scratch[STATIC_SCRATCH_WORDS];

// The array's device address on each device (host code must not pass the
// symbol itself to a kernel: it names a host-side stand-in).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
uint32_t *g_scratch[MAX_DEVICES] = {};

// The address of `scratch` on the current device, resolved once per device
// with that device current.
int scratch_on_current_device(uint32_t **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return ATQ_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return ATQ_ERR_HIP;
        g_scratch[dev] = static_cast<uint32_t *>(ptr);
    }
    *out = g_scratch[dev];
    return ATQ_OK;
}

int check_dims(size_t w, size_t h) {
    if (w < 8 || h < 8 || w > 8192 || h > 8192 || (w & 3) || (h & 3)) return ATQ_ERR_BAD_DIMS;
    return ATQ_OK;
}

int ensure_init() {
    if (g_ready) return ATQ_OK;
    return atq_init();
}

hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int atq_init(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ATQ_ERR_NO_DEVICE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ATQ_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return ATQ_ERR_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return ATQ_ERR_NO_DEVICE;
    uint32_t *scr = nullptr;  // the static segmentation scratch of this device
    if (int r = scratch_on_current_device(&scr)) return r;
    g_ready = true;
    return ATQ_OK;
}

const char *atq_status_string(int status) {
    switch (status) {
    case ATQ_OK: return "ok";
    case ATQ_ERR_BAD_DIMS: return "width/height must be multiples of 4 in [8, 8192]";
    case ATQ_ERR_CAPACITY: return "a count exceeded its buffer capacity";
    case ATQ_ERR_HIP: return "HIP runtime error";
    case ATQ_ERR_BAD_ARG: return "bad argument";
    case ATQ_ERR_NO_DEVICE: return "no gfx950 device";
    default: return "unknown status";
    }
}

int atq_decimate(void *stream, const uint8_t *gray_in, size_t width_in, size_t height_in,
                 int factor, uint8_t *gray_out) {
    if (!gray_in || !gray_out || factor < 1 || width_in == 0 || height_in == 0) return ATQ_ERR_BAD_ARG;
    if (int r = ensure_init()) return r;
    return launch_decimate(S(stream), gray_in, width_in, height_in, factor, gray_out);
}

int atq_grey_decimate(void *stream, const uint8_t *pixels_in, size_t width_in, size_t height_in,
                      int channels, int factor, uint8_t *gray_out) {
    if (!pixels_in || !gray_out || factor < 1 || width_in == 0 || height_in == 0 ||
        (channels != 1 && channels != 3 && channels != 4))
        return ATQ_ERR_BAD_ARG;
    if (int r = ensure_init()) return r;
    return launch_decimate(S(stream), pixels_in, width_in, height_in, factor, gray_out, channels);
}

int atq_threshold_image(void *stream, const uint8_t *grayscale, atq_extrema *extrema,
                        atq_extrema *filtered_extrema, uint8_t *thresholded, size_t width,
                        size_t height) {
    if (!grayscale || !thresholded) return ATQ_ERR_BAD_ARG;
    if (int r = check_dims(width, height)) return r;
    if (int r = ensure_init()) return r;
    return launch_threshold(S(stream), grayscale, width, height, 1, nullptr, extrema,
                            filtered_extrema, thresholded, 5);
}

int atq_decimate_threshold(void *stream, const uint8_t *gray_in, size_t width_in,
                           size_t height_in, int factor, uint8_t *gray_out, atq_extrema *extrema,
                           atq_extrema *filtered_extrema, uint8_t *thresholded) {
    if (!gray_in || !thresholded || factor < 1 || width_in == 0 || height_in == 0)
        return ATQ_ERR_BAD_ARG;
    if (factor > 1 && !gray_out) return ATQ_ERR_BAD_ARG;
    size_t w = factor > 1 ? 1 + (width_in - 1) / factor : width_in;
    size_t h = factor > 1 ? 1 + (height_in - 1) / factor : height_in;
    if (int r = check_dims(w, h)) return r;
    if (int r = ensure_init()) return r;
    return launch_threshold(S(stream), gray_in, width_in, height_in, factor, gray_out, extrema,
                            filtered_extrema, thresholded, 5);
}

int atq_image_segmentation(void *stream, const uint8_t *thresholded, uint32_t *label_scratch,
                           uint32_t *labels, uint32_t *sizes, size_t sizes_elem, size_t width,
                           size_t height) {
    (void)sizes_elem;  // unused, as in the reference (src/segmentation.cpp:97)
    if (!thresholded || !label_scratch || !labels || !sizes) return ATQ_ERR_BAD_ARG;
    if (int r = check_dims(width, height)) return r;
    if (int r = ensure_init()) return r;
    return launch_ccl(S(stream), thresholded, width, height, label_scratch, labels, sizes);
}

int atq_image_segmentation_static(void *stream, const uint8_t *thresholded, uint32_t *labels,
                                  uint32_t *sizes, size_t sizes_elem, size_t width,
                                  size_t height) {
    (void)sizes_elem;
    if (!thresholded || !labels || !sizes) return ATQ_ERR_BAD_ARG;
    if (int r = check_dims(width, height)) return r;
    if (width * height > ATQ_SEGMENTATION_STATIC_MAX_PIXELS) return ATQ_ERR_CAPACITY;
    if (int r = ensure_init()) return r;
    uint32_t *scr = nullptr;
    if (int r = scratch_on_current_device(&scr)) return r;
    return launch_ccl(S(stream), thresholded, width, height, scr, labels, sizes);
}

int atq_find_boundaries(void *stream, const uint32_t *labels, const uint32_t *sizes,
                        atq_boundary_point *points, uint64_t *blob_labels, size_t width,
                        size_t height) {
    if (!labels || !sizes || !points || !blob_labels) return ATQ_ERR_BAD_ARG;
    if (int r = check_dims(width, height)) return r;
    if (int r = ensure_init()) return r;
    return launch_find_boundaries_sparse(S(stream), labels, sizes, width, height, points,
                                         blob_labels, 25);
}

static_assert(sizeof(atq_tag_layout) == 148 && sizeof(atq_tag_code) == 48 &&
                  sizeof(atq_tag_match) == 16,
              "decode records as pyatq/records.py lays them out");

int atq_decode_payloads(void *stream, const uint8_t *gray, size_t width, size_t height,
                        const atq_quad_corners *quads, size_t n_quads, const uint32_t *n_quads_dev,
                        const atq_tag_layout *layout, atq_tag_code *out) {
    if (!gray || !layout || ((!quads || !out) && (n_quads || n_quads_dev))) return ATQ_ERR_BAD_ARG;
    if (width < 1 || height < 1 || width > 0x7FFFFFFFu || height > 0x7FFFFFFFu) return ATQ_ERR_BAD_DIMS;
    if (n_quads > 0xFFFFFFFFu) return ATQ_ERR_BAD_ARG;
    if (int r = ensure_init()) return r;
    return launch_decode_payloads(S(stream), gray, width, height, quads, n_quads, n_quads_dev,
                                  *layout, out);
}

int atq_match_codes(void *stream, const atq_tag_code *codes, size_t n_codes,
                    const uint32_t *n_codes_dev, const uint64_t *table, size_t n_table,
                    const atq_tag_layout *layout, int max_hamming, atq_tag_match *out) {
    if (!layout || ((!codes || !out) && (n_codes || n_codes_dev)) || (!table && n_table))
        return ATQ_ERR_BAD_ARG;
    if (n_codes > 0xFFFFFFFFu) return ATQ_ERR_BAD_ARG;
    if (int r = ensure_init()) return r;
    return launch_match_codes(S(stream), codes, n_codes, n_codes_dev, table, n_table, *layout,
                              max_hamming, out);
}

int atq_linefit_corners_forward(void *stream, const float *image, size_t width, size_t height,
                                const atq_cluster_point *points, size_t n_points,
                                const float *boundary_weights, const atq_diff_quad *quads,
                                size_t n_quads, float *corners, float *w_out) {
    if (!image || !points || !quads || !corners) return ATQ_ERR_BAD_ARG;
    if (width < 3 || height < 3 || n_points > 0xFFFFFFFFu || n_quads > 0xFFFFFFFFu)
        return ATQ_ERR_BAD_DIMS;
    if (int r = ensure_init()) return r;
    return launch_diff_forward(S(stream), image, width, height, points, boundary_weights, quads,
                               n_quads, corners, w_out);
}

int atq_linefit_corners_backward(void *stream, const float *image, size_t width, size_t height,
                                 const atq_cluster_point *points, size_t n_points,
                                 const float *boundary_weights, const atq_diff_quad *quads,
                                 size_t n_quads, const float *grad_corners, float *grad_weights,
                                 float *grad_image) {
    if (!image || !points || !quads || !grad_corners) return ATQ_ERR_BAD_ARG;
    if (width < 3 || height < 3 || n_points > 0xFFFFFFFFu || n_quads > 0xFFFFFFFFu)
        return ATQ_ERR_BAD_DIMS;
    if (int r = ensure_init()) return r;
    return launch_diff_backward(S(stream), image, width, height, points, boundary_weights, quads,
                                n_quads, grad_corners, grad_weights, grad_image);
}

}  // extern "C"
