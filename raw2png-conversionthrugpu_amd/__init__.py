"""MI355X (gfx950) PNG encoder -- Python bindings over libr2p.so's C-ABI.

The product is the C-ABI library (include/r2p.h) and its C++ drop-in facade;
this module is a thin ctypes mirror used by tests, the bench and scripting.
It never falls back to a CPU path: if libr2p.so is missing or no device is
present, calls raise.

Device memory is passed as raw device pointers (ints), e.g. a torch CUDA
tensor's ``data_ptr()``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libr2p.so")
_lib = None

R2P_OK, R2P_ERR_ARG, R2P_ERR_HIP, R2P_ERR_SIZE, R2P_ERR_NODEV, R2P_ERR_UNSUPPORTED, R2P_ERR_STATE = range(7)
PIX_READY, PIX_SWAP16, PIX_DICOM, PIX_DICOM8 = range(4)


class R2PError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"r2p error {code}: {msg}")
        self.code = code


class DicomParams(C.Structure):
    """DicomPixelParams layout (include/dicom_params.h:6-21)."""

    _fields_ = [
        ("bits_allocated", C.c_int), ("bits_stored", C.c_int), ("high_bit", C.c_int),
        ("pixel_rep", C.c_int), ("apply_rescale", C.c_int), ("rescale_slope", C.c_float),
        ("rescale_intercept", C.c_float), ("apply_window", C.c_int),
        ("window_center", C.c_float), ("window_width", C.c_float),
    ]

    @classmethod
    def make(cls, bits_allocated=16, bits_stored=16, high_bit=None, pixel_rep=0,
             apply_rescale=0, rescale_slope=1.0, rescale_intercept=0.0,
             apply_window=0, window_center=32768.0, window_width=65536.0):
        if high_bit is None:
            high_bit = bits_stored - 1
        return cls(bits_allocated, bits_stored, high_bit, pixel_rep, apply_rescale,
                   rescale_slope, rescale_intercept, apply_window, window_center, window_width)


class Image(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_int),
        ("in_bits", C.c_int), ("pix_mode", C.c_int), ("src_pitch", C.c_int64),
        ("dicom", DicomParams),
    ]

    @classmethod
    def make(cls, width, height, channels=1, in_bits=8, pix_mode=PIX_READY, dicom=None,
             src_pitch=0):
        return cls(width, height, channels, in_bits, pix_mode, src_pitch,
                   dicom if dicom is not None else DicomParams.make())


class Options(C.Structure):
    _fields_ = [
        ("segment_bytes", C.c_int), ("lz77", C.c_int), ("idat_chunk_bytes", C.c_int),
        ("zlib_level", C.c_int), ("band_rows", C.c_int), ("lane_bytes", C.c_int),
        ("wg_threads", C.c_int), ("slice_bytes", C.c_int), ("huffman", C.c_int),
    ]


class Timings(C.Structure):
    _fields_ = [("filter_ms", C.c_float), ("deflate_ms", C.c_float), ("scan_ms", C.c_float),
                ("assemble_ms", C.c_float), ("total_ms", C.c_float)]


class LzStats(C.Structure):
    _fields_ = [("matches", C.c_uint64), ("matched_bytes", C.c_uint64), ("positions", C.c_uint64),
                ("max_len", C.c_uint32), ("avg_len", C.c_double), ("avg_dist", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_batch_images", C.c_int),
                ("inflight_batches", C.c_int), ("opt", Options)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("mean_batch_images", C.c_double), ("peak_admitted_bytes", C.c_uint64)]


def lib():
    """Loads libr2p.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, sz, i64, u32, u64 = C.c_void_p, C.c_size_t, C.c_int64, C.c_uint32, C.c_uint64
        ip, imgp, optp = C.c_int, C.POINTER(Image), C.POINTER(Options)
        sig = {
            "r2p_last_error": ([], C.c_char_p), "r2p_version": ([], C.c_char_p),
            "r2p_init": ([ip], ip),
            "r2p_image_out_bits": ([imgp], ip), "r2p_image_row_bytes": ([imgp], sz),
            "r2p_image_filtered_bytes": ([imgp], sz), "r2p_image_input_bytes": ([imgp], sz),
            "r2p_png_bound": ([imgp, ip, ip], sz),
            "r2p_options_default": ([optp], None),
            "r2p_encoder_create": ([sz, optp, C.POINTER(vp)], ip),
            "r2p_encoder_destroy": ([vp], ip), "r2p_encoder_stream": ([vp], vp),
            "r2p_encode_device": ([vp, imgp, vp, vp, sz, vp], ip),
            "r2p_encode_device_noalloc": ([imgp, vp, vp, sz, vp, optp, vp], ip),
            "r2p_encoder_sync": ([vp, C.POINTER(sz), C.POINTER(u32), C.POINTER(u32)], ip),
            "r2p_encode_host": ([vp, imgp, vp, vp, sz, C.POINTER(sz)], ip),
            "r2p_encoder_set_timing": ([vp, ip], ip),
            "r2p_encoder_timings": ([vp, C.POINTER(Timings)], ip),
            "r2p_filter_rows": ([imgp, vp, ip, vp, vp, vp, ip, vp], ip),
            "r2p_checksum_work_bytes": ([sz], sz),
            "r2p_adler32_device": ([vp, sz, vp, vp, vp], ip),
            "r2p_crc32_device": ([vp, sz, vp, vp, vp], ip),
            "r2p_adler32_combine": ([u32, u32, u64], u32),
            "r2p_crc32_combine": ([u32, u32, u64], u32),
            "r2p_deflate_create": ([sz, sz, ip, ip, C.POINTER(vp)], ip),
            "r2p_deflate_destroy": ([vp], ip), "r2p_deflate_reset": ([vp], ip),
            "r2p_deflate_strip": ([vp, vp, sz, ip, ip, ip], ip),
            "r2p_deflate_sync": ([vp], ip), "r2p_deflate_output": ([vp], vp),
            "r2p_deflate_output_bytes": ([vp], sz), "r2p_deflate_adler": ([vp], u32),
            "r2p_deflate_copy_to_host": ([vp, vp, sz], ip), "r2p_deflate_stream": ([vp], vp),
            "r2p_deflate_lz_stats": ([vp, C.POINTER(LzStats)], ip),
            "r2p_deflate_rewind": ([vp], ip),
            "r2p_idat_chunk": ([vp, sz, ip, ip, u32, ip, vp, sz, vp, C.POINTER(sz), vp], ip),
            "r2p_batch_create": ([ip, sz, sz, optp, C.POINTER(vp)], ip),
            "r2p_batch_destroy": ([vp], ip),
            "r2p_batch_set": ([vp, ip, imgp, C.POINTER(vp)], ip),
            "r2p_batch_run": ([vp, ip], ip), "r2p_batch_sync": ([vp], ip),
            "r2p_batch_result": ([vp, ip, C.POINTER(vp), C.POINTER(sz)], ip),
            "r2p_batch_copy_result": ([vp, ip, vp, sz, C.POINTER(sz)], ip),
            "r2p_batch_timings": ([vp, C.POINTER(Timings)], ip),
            "r2p_batch_autotune": ([vp, optp, C.POINTER(C.c_double)], ip),
            "r2p_batch_options": ([vp, optp], ip),
            "r2p_batch_job_options": ([vp, ip, optp], ip),
            "r2p_batch_class_slices": ([vp, C.POINTER(C.c_int)], ip),
            "r2p_server_create": ([C.POINTER(ServerConfig), C.POINTER(vp)], ip),
            "r2p_server_destroy": ([vp], ip),
            "r2p_server_submit": ([vp, imgp, vp, C.POINTER(i64)], ip),
            "r2p_server_wait": ([vp, i64, ip, vp, sz, C.POINTER(sz)], ip),
            "r2p_server_stats_get": ([vp, C.POINTER(ServerStats)], ip),
            "r2p_server_loadtest": ([vp, ip, C.POINTER(C.c_double), C.POINTER(ip), imgp,
                                     C.POINTER(vp), C.POINTER(C.c_double), C.POINTER(u64),
                                     C.POINTER(C.c_double), C.POINTER(vp), C.POINTER(sz),
                                     C.POINTER(C.c_uint8)], ip),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != R2P_OK:
        raise R2PError(rc, lib().r2p_last_error().decode())
    return rc


def options(**kw) -> Options:
    o = Options()
    lib().r2p_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def filtered_bytes(img: Image) -> int:
    return lib().r2p_image_filtered_bytes(C.byref(img))


def row_bytes(img: Image) -> int:
    return lib().r2p_image_row_bytes(C.byref(img))


def input_bytes(img: Image) -> int:
    return lib().r2p_image_input_bytes(C.byref(img))


def out_bits(img: Image) -> int:
    return lib().r2p_image_out_bits(C.byref(img))


def png_bound(img: Image, opt: Options | None = None) -> int:
    opt = opt or options()
    return lib().r2p_png_bound(C.byref(img), opt.segment_bytes, opt.idat_chunk_bytes)


class Encoder:
    """Plain per-image device encoder (r2p_encoder_*)."""

    def __init__(self, max_filtered_bytes: int, opt: Options | None = None, timing: bool = False):
        self.opt = opt or options()
        h = C.c_void_p()
        check(lib().r2p_encoder_create(max_filtered_bytes, C.byref(self.opt), C.byref(h)))
        self.h = h
        if timing:
            self.set_timing(True)

    def set_timing(self, on: bool):
        """Per-kernel HIP events on (timings()) or off (the default: each event
        costs the stream about 3 us per encode)."""
        check(lib().r2p_encoder_set_timing(self.h, 1 if on else 0))

    def close(self):
        if self.h:
            lib().r2p_encoder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return lib().r2p_encoder_stream(self.h) or 0

    def encode_device(self, img: Image, d_src: int, d_png: int, png_cap: int, d_filtered: int = 0):
        check(lib().r2p_encode_device(self.h, C.byref(img), d_src, d_png, png_cap, d_filtered or None))

    def sync(self):
        n, a, s = C.c_size_t(), C.c_uint32(), C.c_uint32()
        check(lib().r2p_encoder_sync(self.h, C.byref(n), C.byref(a), C.byref(s)))
        return n.value, a.value, s.value

    def encode_host(self, img: Image, pixels: np.ndarray) -> bytes:
        pixels = np.ascontiguousarray(pixels)
        cap = png_bound(img, self.opt)
        out = np.empty(cap, np.uint8)
        n = C.c_size_t()
        check(lib().r2p_encode_host(self.h, C.byref(img), pixels.ctypes.data, out.ctypes.data, cap,
                                    C.byref(n)))
        return out[: n.value].tobytes()

    def timings(self) -> Timings:
        t = Timings()
        check(lib().r2p_encoder_timings(self.h, C.byref(t)))
        return t


def filter_rows(img: Image, d_src: int, rows: int, d_filtered: int, d_prior: int = 0,
                d_last_row: int = 0, band_rows: int = 0, stream: int = 0):
    check(lib().r2p_filter_rows(C.byref(img), d_src, rows, d_prior or None, d_filtered,
                                d_last_row or None, band_rows, stream or None))


def checksum_work_bytes(n: int) -> int:
    return lib().r2p_checksum_work_bytes(n)


def adler32_device(d_data: int, n: int, d_work: int, d_result: int, stream: int = 0):
    check(lib().r2p_adler32_device(d_data or None, n, d_work, d_result, stream or None))


def crc32_device(d_data: int, n: int, d_work: int, d_result: int, stream: int = 0):
    check(lib().r2p_crc32_device(d_data or None, n, d_work, d_result, stream or None))


class StripDeflate:
    """Strip-incremental device DEFLATE (r2p_deflate_*)."""

    def __init__(self, max_strip: int, max_total: int, lz77: int = 1, segment_bytes: int = 65536):
        h = C.c_void_p()
        check(lib().r2p_deflate_create(max_strip, max_total, lz77, segment_bytes, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().r2p_deflate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(lib().r2p_deflate_reset(self.h))

    def strip(self, d_filtered: int, n: int, is_last: bool, bpp: int = 1, row_stride: int = 0):
        check(lib().r2p_deflate_strip(self.h, d_filtered or None, n, int(is_last), bpp, row_stride))

    def sync(self):
        check(lib().r2p_deflate_sync(self.h))

    def output(self) -> bytes:
        self.sync()
        n = lib().r2p_deflate_output_bytes(self.h)
        out = np.empty(max(n, 1), np.uint8)
        check(lib().r2p_deflate_copy_to_host(self.h, out.ctypes.data, n))
        return out[:n].tobytes()

    def adler(self) -> int:
        return lib().r2p_deflate_adler(self.h)

    def lz_stats(self) -> LzStats:
        s = LzStats()
        check(lib().r2p_deflate_lz_stats(self.h, C.byref(s)))
        return s


def idat_chunk(d_slice: int, n: int, is_first: bool, is_last: bool, adler: int, level: int,
               d_chunk: int, cap: int, d_work: int, stream: int = 0) -> int:
    out = C.c_size_t()
    check(lib().r2p_idat_chunk(d_slice or None, n, int(is_first), int(is_last), adler, level, d_chunk,
                               cap, d_work, C.byref(out), stream or None))
    return out.value


class Batch:
    """Batched form (r2p_batch_*): many images per launch sequence / HIP graph."""

    def __init__(self, max_images: int, max_input: int, max_filtered: int, opt: Options | None = None):
        self.opt = opt or options()
        h = C.c_void_p()
        check(lib().r2p_batch_create(max_images, max_input, max_filtered, C.byref(self.opt), C.byref(h)))
        self.h = h
        self.n = 0
        self._keep = None

    def close(self):
        if self.h:
            lib().r2p_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, imgs, d_srcs):
        n = len(imgs)
        arr = (Image * n)(*imgs)
        ptrs = (C.c_void_p * n)(*[C.c_void_p(p) for p in d_srcs])
        self._keep = (arr, ptrs)
        check(lib().r2p_batch_set(self.h, n, arr, ptrs))
        self.n = n

    def set_options(self, opt: Options):
        check(lib().r2p_batch_options(self.h, C.byref(opt)))
        self.opt = opt

    def run(self, graph: bool = True):
        check(lib().r2p_batch_run(self.h, int(graph)))

    def sync(self):
        check(lib().r2p_batch_sync(self.h))

    def result(self, i: int):
        p, n = C.c_void_p(), C.c_size_t()
        check(lib().r2p_batch_result(self.h, i, C.byref(p), C.byref(n)))
        return p.value, n.value

    def copy_result(self, i: int) -> bytes:
        _, n = self.result(i)
        out = np.empty(max(n, 1), np.uint8)
        m = C.c_size_t()
        check(lib().r2p_batch_copy_result(self.h, i, out.ctypes.data, out.size, C.byref(m)))
        return out[: m.value].tobytes()

    def timings(self) -> Timings:
        t = Timings()
        check(lib().r2p_batch_timings(self.h, C.byref(t)))
        return t

    def autotune(self):
        o, ms = Options(), C.c_double()
        check(lib().r2p_batch_autotune(self.h, C.byref(o), C.byref(ms)))
        self.opt = o
        return o, ms.value

    def job_options(self, i: int) -> Options:
        """The options image i is encoded with (its slice size resolved): the
        plain encoder with them produces the same PNG."""
        o = Options()
        check(lib().r2p_batch_job_options(self.h, i, C.byref(o)))
        return o

    def class_slices(self):
        v = (C.c_int * 3)()
        check(lib().r2p_batch_class_slices(self.h, v))
        return list(v)


def server_loadtest(server, arrival_s, kind, imgs, h_srcs, expect_pngs=None):
    """Open-loop load test through r2p_server_loadtest.  arrival_s: seconds
    after start; kind: index into imgs / h_srcs (host pointers) per request;
    expect_pngs: optional bytes per kind that every request's PNG must equal.
    Returns (latency_s, png_len, wall_s, same) -- same is None without
    expected PNGs, else a uint8 array (1 where the bytes matched)."""
    n = len(arrival_s)
    arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
    kd = np.ascontiguousarray(kind, dtype=np.int32)
    im = (Image * len(imgs))(*imgs)
    src = (C.c_void_p * len(h_srcs))(*[C.c_void_p(p) for p in h_srcs])
    lat = np.zeros(n)
    plen = np.zeros(n, np.uint64)
    wall = C.c_double()
    same = ex = exl = None
    keep = None
    if expect_pngs is not None:
        keep = [np.frombuffer(b, np.uint8) for b in expect_pngs]
        ex = (C.c_void_p * len(keep))(*[C.c_void_p(k.ctypes.data) for k in keep])
        exl = (C.c_size_t * len(keep))(*[k.size for k in keep])
        same = np.zeros(n, np.uint8)
    check(lib().r2p_server_loadtest(server, n, arr.ctypes.data_as(C.POINTER(C.c_double)),
                                    kd.ctypes.data_as(C.POINTER(C.c_int)), im, src,
                                    lat.ctypes.data_as(C.POINTER(C.c_double)),
                                    plen.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(wall), ex, exl,
                                    same.ctypes.data_as(C.POINTER(C.c_uint8)) if same is not None else None))
    del keep
    return lat, plen, wall.value, same
