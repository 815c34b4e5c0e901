"""ctypes binding of libatq.so (include/atq.h) for tests and the benchmark.

Device memory, streams and events come straight from the HIP runtime that
libatq.so links (no torch: its bundled HIP runtime shares the soname and
must not be mixed in); every compute call goes through the C ABI into the
HIP kernels. There is no fallback: if
libatq.so is missing or no gfx950 device is visible, calls raise AtqError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import records as R

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libatq.so")
INCLUDE_H = os.path.join(os.path.dirname(PKG_DIR), "include", "atq.h")

_P = C.c_void_p
_S = C.c_size_t

# name -> (restype, argtypes); mirrors include/atq.h
SIGNATURES = {
    "atq_init": (C.c_int, []),
    "atq_status_string": (C.c_char_p, [C.c_int]),
    "atq_decimate": (C.c_int, [_P, _P, _S, _S, C.c_int, _P]),
    "atq_grey_decimate": (C.c_int, [_P, _P, _S, _S, C.c_int, C.c_int, _P]),
    "atq_threshold_image": (C.c_int, [_P, _P, _P, _P, _P, _S, _S]),
    "atq_decimate_threshold": (C.c_int, [_P, _P, _S, _S, C.c_int, _P, _P, _P, _P]),
    "atq_image_segmentation": (C.c_int, [_P, _P, _P, _P, _P, _S, _S, _S]),
    "atq_image_segmentation_static": (C.c_int, [_P, _P, _P, _P, _S, _S, _S]),
    "atq_find_boundaries": (C.c_int, [_P, _P, _P, _P, _P, _S, _S]),
    "atq_glue_workspace_size": (_S, [_S, _S]),
    "atq_group_boundaries": (C.c_int, [_P, _P, _P, _S, _S, _P, _P, _P, _P, _S]),
    "atq_cluster_bounds_filter": (C.c_int, [_P, _P, _P, _P, _S, _S, _P, _P, _P, _S]),
    "atq_cluster_points_sort": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _S]),
    "atq_linefit_prefix": (C.c_int, [_P, _P, _P, _P, _P, _S, _S, _P]),
    "atq_fit_lines": (C.c_int, [_P, _P, _P, _P, _S, _P]),
    "atq_corner_segments": (C.c_int, [_P, _P, _P, _P, _P, _P, _S]),
    "atq_do_indexing": (C.c_int, [_P, _P, _S, _P, _P, _S, _P, _P]),
    "atq_quads_from_fits": (C.c_int, [_P, _P, _S, _P, _P, _P, _S]),
    "atq_detector_create": (C.c_int, [_S, _S, _P, C.POINTER(_P)]),
    "atq_detector_destroy": (C.c_int, [_P]),
    "atq_detect_quads": (C.c_int, [_P, _P, _P, _S, _S]),
    "atq_detect_quads_rgb": (C.c_int, [_P, _P, _P, _S, _S, C.c_int]),
    "atq_detector_buffer": (C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(_S)]),
    "atq_detector_dims": (C.c_int, [_P, C.POINTER(_S), C.POINTER(_S)]),
    "atq_detector_set_timing": (C.c_int, [_P, C.c_int]),
    "atq_detector_stage_times": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int]),
    "atq_stage_name": (C.c_char_p, [C.c_int]),
    "atq_batch_create": (C.c_int, [_S, C.POINTER(_S), C.POINTER(_S), C.POINTER(C.c_int32),
                                   C.POINTER(_P)]),
    "atq_batch_destroy": (C.c_int, [_P]),
    "atq_batch_input": (C.c_int, [_P, _S, C.POINTER(_P)]),
    "atq_batch_detect": (C.c_int, [_P, _P, C.c_int]),
    "atq_batch_frame": (C.c_int, [_P, _S, C.POINTER(_P)]),
    "atq_batch_autotune": (C.c_int, [_P, _P]),
    "atq_detector_set_shapes": (C.c_int, [_P, _P]),
    "atq_detector_get_shapes": (C.c_int, [_P, _P]),
    "atq_autotune": (C.c_int, [_P, _P, _P, _S, _S, _P]),
    "atq_decode_payloads": (C.c_int, [_P, _P, _S, _S, _P, _S, _P, _P, _P]),
    "atq_match_codes": (C.c_int, [_P, _P, _S, _P, _P, _S, _P, C.c_int, _P]),
    "atq_linefit_corners_forward": (C.c_int, [_P, _P, _S, _S, _P, _S, _P, _P, _S, _P, _P]),
    "atq_linefit_corners_backward": (C.c_int, [_P, _P, _S, _S, _P, _S, _P, _P, _S, _P, _P, _P]),
}
N_STAGES = 10

BUF = dict(GRAY=0, THRESHOLDED=1, LABELS=2, SIZES=3, BOUNDARY_POINTS=4, BOUNDARY_KEYS=5,
           CLUSTER_BOUNDS=6, CLUSTER_POINTS=7, CLUSTER_INDEX=8, CLUSTER_EXTENTS=9, LINE_FIT=10,
           CORNERS=11, PEAK_EXTENTS=12, FITTED=13, QUADS=14, COUNTS=15, EXTREMA=16,
           FILTERED_EXTREMA=17, SPARSE_CORNERS=18, PAIR_KEYS=19, PAIR_COUNTS=20)


class AtqError(RuntimeError):
    """A non-zero atq status; .code is the ATQ_ERR_* value (0 when the error
    comes from the binding itself)."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class AtqParams(C.Structure):
    _fields_ = [("decimate", C.c_int32), ("min_white_black_diff", C.c_int32),
                ("min_component_size", C.c_int32), ("min_tag_width", C.c_int32),
                ("pair_table_log2", C.c_int32), ("mode", C.c_int32)]


class AtqShapes(C.Structure):
    _fields_ = [("threshold_tiles_x", C.c_int32), ("threshold_tiles_y", C.c_int32),
                ("ccl_block_rows", C.c_int32)]

    def as_tuple(self):
        return (self.threshold_tiles_x, self.threshold_tiles_y, self.ccl_block_rows)


# every (threshold_tiles_x, threshold_tiles_y, ccl_block_rows) the library takes
SHAPES = [(rx, ry, rows) for rows in (16, 8, 32) for rx, ry in ((64, 16), (32, 32), (16, 64))]

_lib = None


def lib():
    """The loaded libatq.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AtqError(f"{LIB_PATH} is missing: build it with `make -C {PKG_DIR}`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().atq_status_string(status).decode()
        raise AtqError(f"atq status {status}: {msg}", status)


# ---- HIP plumbing (the same libamdhip64 instance libatq.so links) ---------

class Batch:
    """Frames of uneven sizes through one graph launch (atq_batch_*)."""

    def __init__(self, sizes, decimates):
        init()
        n = len(sizes)
        ws = (C.c_size_t * n)(*[w for w, _ in sizes])
        hs = (C.c_size_t * n)(*[h for _, h in sizes])
        ds = (C.c_int32 * n)(*decimates)
        h = C.c_void_p()
        check(lib().atq_batch_create(n, ws, hs, ds, C.byref(h)))
        self.handle, self.sizes = h, list(sizes)

    def upload(self, i, img):
        p = C.c_void_p()
        check(lib().atq_batch_input(self.handle, i, C.byref(p)))
        a = np.ascontiguousarray(img)
        hip_check(_hip().hipMemcpy(p, C.c_void_p(a.ctypes.data), a.nbytes, 1), "hipMemcpy H2D")

    def detect(self, stream_handle, use_graph=True):
        check(lib().atq_batch_detect(self.handle, C.c_void_p(stream_handle), 1 if use_graph else 0))

    def frame(self, i):
        """A Detector view (not owning) of frame i's buffers."""
        p = C.c_void_p()
        check(lib().atq_batch_frame(self.handle, i, C.byref(p)))
        d = Detector.__new__(Detector)
        d.handle = p
        d._borrowed = True
        return d

    def close(self):
        if self.handle:
            lib().atq_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_hiplib = None


def _hip():
    global _hiplib
    if _hiplib is None:
        lib()  # loads libatq.so and, with it, its HIP runtime
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        H.hipDeviceSynchronize.argtypes = []
        H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipStreamDestroy.argtypes = [C.c_void_p]
        H.hipStreamSynchronize.argtypes = [C.c_void_p]
        H.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        H.hipEventDestroy.argtypes = [C.c_void_p]
        H.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        H.hipEventSynchronize.argtypes = [C.c_void_p]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        H.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize",
                  "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize",
                  "hipEventCreate", "hipEventDestroy", "hipEventRecord", "hipEventSynchronize",
                  "hipEventElapsedTime", "hipGetDeviceCount"):
            getattr(H, f).restype = C.c_int
        _hiplib = H
    return _hiplib


def hip_check(rc, what):
    if rc != 0:
        raise AtqError(f"{what} failed with hipError {rc}")


class DevBuf:
    """A hipMalloc'd device buffer (freed with the object)."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        p = C.c_void_p()
        hip_check(_hip().hipMalloc(C.byref(p), self.nbytes), "hipMalloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            _hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def data_ptr(self):
        return self.ptr

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        hip_check(_hip().hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes, 1),
                  "hipMemcpy H2D")
        return self

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize if count is None else int(count)
        out = np.empty(n, dtype)
        if n:
            hip_check(_hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr),
                                       n * dtype.itemsize, 2), "hipMemcpy D2H")
        return out

    def fill(self, byte):
        hip_check(_hip().hipMemset(C.c_void_p(self.ptr), int(byte), self.nbytes), "hipMemset")
        return self


def gpu_available():
    try:
        n = C.c_int(0)
        return _hip().hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except Exception:
        return False


def init():
    check(lib().atq_init())


def dev_empty(nbytes):
    return DevBuf(nbytes)


def dev_from_np(a):
    a = np.ascontiguousarray(a)
    return DevBuf(a.nbytes).upload(a)


def to_np(buf, dtype, count=None, shape=None):
    a = buf.download(dtype, count)
    return a.reshape(shape) if shape is not None else a


def P(buf):
    return C.c_void_p(buf.data_ptr()) if buf is not None else C.c_void_p(0)


_default_stream = None


def stream():
    """The stream every wrapper launches on (NULL = HIP's default stream)."""
    return C.c_void_p(_default_stream)


def sync():
    hip_check(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        hip_check(_hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def sync(self):
        hip_check(_hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def __del__(self):
        try:
            _hip().hipStreamDestroy(C.c_void_p(self.handle))
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        hip_check(_hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream_handle=None):
        hip_check(_hip().hipEventRecord(C.c_void_p(self.handle), C.c_void_p(stream_handle)),
                  "hipEventRecord")

    def synchronize(self):
        hip_check(_hip().hipEventSynchronize(C.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        hip_check(_hip().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.handle),
                                             C.c_void_p(end.handle)), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            _hip().hipEventDestroy(C.c_void_p(self.handle))
        except Exception:
            pass


# ---- stage drop-ins ----------------------------------------------------------

def threshold_image(gray_dev, w, h):
    ext = dev_empty((w // 4) * (h // 4) * 2)
    fil = dev_empty((w // 4) * (h // 4) * 2)
    thr = dev_empty(w * h)
    check(lib().atq_threshold_image(stream(), P(gray_dev), P(ext), P(fil), P(thr), w, h))
    return ext, fil, thr


def decimate_threshold(gray_dev, w_in, h_in, f, want_extrema=True):
    w, h = (w_in, h_in) if f == 1 else (1 + (w_in - 1) // f, 1 + (h_in - 1) // f)
    gray = dev_empty(w * h) if f > 1 else None
    ext = dev_empty((w // 4) * (h // 4) * 2) if want_extrema else None
    fil = dev_empty((w // 4) * (h // 4) * 2) if want_extrema else None
    thr = dev_empty(w * h)
    check(lib().atq_decimate_threshold(stream(), P(gray_dev), w_in, h_in, f, P(gray), P(ext),
                                       P(fil), P(thr)))
    return (w, h), gray, ext, fil, thr


def decimate(gray_dev, w_in, h_in, f):
    w, h = 1 + (w_in - 1) // f, 1 + (h_in - 1) // f
    out = dev_empty(w * h)
    check(lib().atq_decimate(stream(), P(gray_dev), w_in, h_in, f, P(out)))
    return (w, h), out


def grey_decimate(pixels_dev, w_in, h_in, channels, f):
    """Colour removal + decimation of interleaved pixels (atq_grey_decimate)."""
    w, h = 1 + (w_in - 1) // f, 1 + (h_in - 1) // f
    out = dev_empty(w * h)
    check(lib().atq_grey_decimate(stream(), P(pixels_dev), w_in, h_in, channels, f, P(out)))
    return (w, h), out


def image_segmentation(thr_dev, w, h):
    scratch = dev_empty(4 * w * h)
    labels = dev_empty(4 * w * h)
    sizes = dev_empty(4 * w * h)
    check(lib().atq_image_segmentation(stream(), P(thr_dev), P(scratch), P(labels), P(sizes),
                                       1 << 16, w, h))
    return labels, sizes


def find_boundaries(labels_dev, sizes_dev, w, h):
    pts = dev_empty(4 * 4 * w * h)
    keys = dev_empty(8 * 4 * w * h)
    check(lib().atq_find_boundaries(stream(), P(labels_dev), P(sizes_dev), P(pts), P(keys), w, h))
    return pts, keys


def tag_layout(d):
    """atq_tag_layout record from a dict (width_at_border, total_width,
    reversed_border, bit_x, bit_y, decode_sharpening)."""
    from . import records as R
    lay = np.zeros(1, R.TAG_LAYOUT)
    n = len(d["bit_x"])
    lay["width_at_border"] = d["width_at_border"]
    lay["total_width"] = d["total_width"]
    lay["reversed_border"] = d.get("reversed_border", 0)
    lay["nbits"] = n
    lay["decode_sharpening"] = d.get("decode_sharpening", 0.25)
    lay["bit_x"][0, :n] = d["bit_x"]
    lay["bit_y"][0, :n] = d["bit_y"]
    return lay


def decode_payloads(gray_dev, w, h, quads_dev, n, layout, n_dev_ptr=None):
    """atq_decode_payloads -> (device buffer of atq_tag_code, its n slots)."""
    lay = tag_layout(layout) if isinstance(layout, dict) else layout
    out = dev_empty(max(n, 1) * 48)
    nd = C.c_void_p(n_dev_ptr) if n_dev_ptr else C.c_void_p(0)
    check(lib().atq_decode_payloads(stream(), P(gray_dev), w, h, quads_dev if isinstance(quads_dev, C.c_void_p)
                                    else P(quads_dev), n, nd,
                                    C.c_void_p(lay.ctypes.data), P(out)))
    return out


def match_codes(codes_dev, n, table_dev, n_table, layout, max_hamming, n_dev_ptr=None):
    """atq_match_codes -> device buffer of atq_tag_match."""
    lay = tag_layout(layout) if isinstance(layout, dict) else layout
    out = dev_empty(max(n, 1) * 16)
    nd = C.c_void_p(n_dev_ptr) if n_dev_ptr else C.c_void_p(0)
    check(lib().atq_match_codes(stream(), P(codes_dev), n, nd, P(table_dev), n_table,
                                C.c_void_p(lay.ctypes.data), int(max_hamming), P(out)))
    return out


class Detector:
    """Owns the per-frame buffers (allocated once) of the whole pipeline."""

    def __init__(self, max_w_in, max_h_in, decimate=1, pair_table_log2=0, fixed=False):
        """fixed=True: the detector's fixed mode (ATQ_MODE_FIXED)."""
        self.params = AtqParams(decimate, 5, 25, 8, pair_table_log2, 1 if fixed else 0)
        h = C.c_void_p()
        init()
        check(lib().atq_detector_create(max_w_in, max_h_in, C.byref(self.params), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle and not getattr(self, "_borrowed", False):
            lib().atq_detector_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def detect(self, gray_dev, w_in, h_in, stream_handle=None):
        s = C.c_void_p(stream_handle) if stream_handle is not None else stream()
        check(lib().atq_detect_quads(self.handle, s, P(gray_dev), w_in, h_in))

    def detect_rgb(self, pixels_dev, w_in, h_in, channels, stream_handle=None):
        """Detection from interleaved RGB (3) or RGBA (4) pixels."""
        s = C.c_void_p(stream_handle) if stream_handle is not None else stream()
        check(lib().atq_detect_quads_rgb(self.handle, s, P(pixels_dev), w_in, h_in, channels))

    def buffer(self, which):
        ptr = C.c_void_p()
        cap = C.c_size_t()
        check(lib().atq_detector_buffer(self.handle, BUF[which], C.byref(ptr), C.byref(cap)))
        return ptr.value, cap.value

    def set_timing(self, on=True):
        check(lib().atq_detector_set_timing(self.handle, 1 if on else 0))

    def set_shapes(self, shapes):
        sh = AtqShapes(*shapes)
        check(lib().atq_detector_set_shapes(self.handle, C.byref(sh)))

    def shapes(self):
        sh = AtqShapes()
        check(lib().atq_detector_get_shapes(self.handle, C.byref(sh)))
        return sh.as_tuple()

    def autotune(self, gray_dev, w_in, h_in, stream_handle=None):
        """Times every shape on this frame and keeps the fastest (atq_autotune)."""
        s = C.c_void_p(stream_handle) if stream_handle is not None else stream()
        sh = AtqShapes()
        check(lib().atq_autotune(self.handle, s, P(gray_dev), w_in, h_in, C.byref(sh)))
        return sh.as_tuple()

    def stage_times(self):
        """Milliseconds per stage of the last frame (needs set_timing(True))."""
        arr = (C.c_float * N_STAGES)()
        check(lib().atq_detector_stage_times(self.handle, arr, N_STAGES))
        return {lib().atq_stage_name(i).decode(): arr[i] for i in range(N_STAGES)}

    def dims(self):
        w, h = C.c_size_t(), C.c_size_t()
        check(lib().atq_detector_dims(self.handle, C.byref(w), C.byref(h)))
        return w.value, h.value

    def read(self, which, dtype, count):
        """Copies `count` elements of a detector buffer to a numpy array."""
        ptr, cap = self.buffer(which)
        dtype = np.dtype(dtype)
        count = int(min(count, cap))
        out = np.empty(count, dtype)
        if count:
            sync()
            hip_check(_hip().hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr),
                                       count * dtype.itemsize, 2), "hipMemcpy D2H")
        return out

    def counts(self):
        return self.read("COUNTS", R.COUNTS, 1)[0]

    def quads(self):
        """The last frame's quads; raises AtqError if the frame failed
        (counts.overflow != 0: a buffer overflowed or a bounded wait of a
        look-back scan or grid barrier gave up, and the results are void)."""
        c = self.counts()
        if int(c["overflow"]):
            raise AtqError(f"frame failed: atq_counts.overflow = {int(c['overflow']):#x}")
        return self.read("QUADS", R.QUAD_CORNERS, int(c["n_quads"]))


