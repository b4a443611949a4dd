"""CPU oracle for the PNG encode hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg import this package, and only as the checker (or the timed CPU baseline).
The product never imports it.

ctypes bindings over ``liboracle.so`` (built from ``r2p_oracle.c`` by
``oracle/Makefile``) plus numpy conveniences.  ``ref()`` returns the bindings
over ``_ref/libref.so`` -- the reference's own zlib-framing sources -- when that
library was built (it needs the reference tree at build time).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None


class DicomParams(C.Structure):
    """Field order and types of DicomPixelParams (include/dicom_params.h:6-21)."""

    _fields_ = [
        ("bits_allocated", C.c_int),
        ("bits_stored", C.c_int),
        ("high_bit", C.c_int),
        ("pixel_rep", C.c_int),
        ("apply_rescale", C.c_int),
        ("rescale_slope", C.c_float),
        ("rescale_intercept", C.c_float),
        ("apply_window", C.c_int),
        ("window_center", C.c_float),
        ("window_width", C.c_float),
    ]

    @classmethod
    def make(cls, bits_allocated=16, bits_stored=16, high_bit=None, pixel_rep=0,
             apply_rescale=0, rescale_slope=1.0, rescale_intercept=0.0,
             apply_window=0, window_center=32768.0, window_width=65536.0):
        if high_bit is None:
            high_bit = bits_stored - 1
        return cls(bits_allocated, bits_stored, high_bit, pixel_rep, apply_rescale,
                   rescale_slope, rescale_intercept, apply_window, window_center,
                   window_width)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _u8p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        u8p, sz = C.POINTER(C.c_uint8), C.c_size_t
        dp = C.POINTER(DicomParams)
        L.r2po_dicom_window.argtypes = [u8p, u8p, sz, dp]
        L.r2po_q15_coeffs.argtypes = [dp, C.POINTER(C.c_int16), C.POINTER(C.c_int32)]
        L.r2po_q15_coeffs.restype = C.c_int
        L.r2po_q15_apply.argtypes = [u8p, u8p, sz, C.c_int16, C.c_int32, C.c_int]
        L.r2po_lut8_build.argtypes = [u8p, dp, C.c_int]
        L.r2po_lut8_apply.argtypes = [u8p, u8p, sz, u8p, C.c_int, C.c_uint16]
        L.r2po_window8.argtypes = [u8p, u8p, sz, dp]
        L.r2po_filter_rows.argtypes = [u8p, u8p, u8p, C.c_int, C.c_int, C.c_int]
        L.r2po_filter_scores.argtypes = [u8p, u8p, C.POINTER(C.c_uint32), C.c_int, C.c_int, C.c_int]
        L.r2po_unfilter_rows.argtypes = [u8p, u8p, u8p, C.c_int, C.c_int, C.c_int]
        L.r2po_unfilter_rows.restype = C.c_int
        L.r2po_adler32.argtypes = [C.c_uint32, u8p, sz]
        L.r2po_adler32.restype = C.c_uint32
        L.r2po_adler32_combine.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        L.r2po_adler32_combine.restype = C.c_uint32
        L.r2po_crc32.argtypes = [C.c_uint32, u8p, sz]
        L.r2po_crc32.restype = C.c_uint32
        L.r2po_crc32_combine.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        L.r2po_crc32_combine.restype = C.c_uint32
        L.r2po_zlib_flg.argtypes = [C.c_int]
        L.r2po_zlib_flg.restype = C.c_uint8
        L.r2po_encode_png.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, dp,
                                      C.c_int, C.c_int, C.c_int, u8p, sz, u8p]
        L.r2po_encode_png.restype = sz
        L.r2po_filtered_size.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.r2po_filtered_size.restype = sz
        L.r2po_zlib_version.restype = C.c_char_p
        _lib = L
    return _lib


def ref():
    """Bindings over oracle/_ref/libref.so, or None when it was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        u8p = C.POINTER(C.c_uint8)
        R.ref_zlib_header.argtypes = [C.c_int, u8p]
        R.ref_zlib_trailer.argtypes = [C.c_uint32, u8p]
        R.ref_deflate_strip.argtypes = [u8p, C.c_size_t, C.c_int, C.c_int, C.c_int, u8p,
                                        C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        R.ref_deflate_strip.restype = C.c_size_t
        R.ref_accum_adler.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_int]
        R.ref_accum_adler.restype = C.c_uint32
        R.ref_png_write.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, u8p,
                                    C.c_size_t]
        R.ref_png_write.restype = C.c_int
        _ref = R
    return _ref


# --------------------------------------------------------------------------
# numpy conveniences
# --------------------------------------------------------------------------
def dicom_window(raw: np.ndarray, p: DicomParams) -> np.ndarray:
    """a2 float window: raw LE samples (uint16 or uint8 array) -> PNG-ready bytes."""
    raw = np.ascontiguousarray(raw)
    n = raw.size
    out = np.empty(n * (2 if p.bits_allocated == 16 else 1), np.uint8)
    lib().r2po_dicom_window(_u8p(raw.view(np.uint8)), _u8p(out), n, C.byref(p))
    return out


def q15_coeffs(p: DicomParams):
    a, b = C.c_int16(), C.c_int32()
    ok = lib().r2po_q15_coeffs(C.byref(p), C.byref(a), C.byref(b))
    return (a.value, b.value) if ok else None


def window8(raw: np.ndarray, p: DicomParams) -> np.ndarray:
    """a5: 16-bit DICOM -> 8-bit (Q15 when it applies, else the LUT)."""
    raw = np.ascontiguousarray(raw, dtype="<u2")
    out = np.empty(raw.size, np.uint8)
    lib().r2po_window8(_u8p(raw.view(np.uint8)), _u8p(out), raw.size, C.byref(p))
    return out


def lut8(p: DicomParams) -> np.ndarray:
    n = 1 << p.bits_stored
    out = np.empty(n, np.uint8)
    lib().r2po_lut8_build(_u8p(out), C.byref(p), n)
    return out


def filter_rows(ready: np.ndarray, rows: int, row_bytes: int, bpp: int,
                prior: np.ndarray | None = None) -> np.ndarray:
    ready = np.ascontiguousarray(ready, dtype=np.uint8).reshape(-1)
    assert ready.size == rows * row_bytes
    out = np.empty(rows * (row_bytes + 1), np.uint8)
    pr = None if prior is None else _u8p(np.ascontiguousarray(prior, dtype=np.uint8))
    lib().r2po_filter_rows(_u8p(ready), pr, _u8p(out), rows, row_bytes, bpp)
    return out


def filter_scores(ready: np.ndarray, rows: int, row_bytes: int, bpp: int) -> np.ndarray:
    ready = np.ascontiguousarray(ready, dtype=np.uint8).reshape(-1)
    out = np.empty(rows * 5, np.uint32)
    lib().r2po_filter_scores(_u8p(ready), None, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                             rows, row_bytes, bpp)
    return out.reshape(rows, 5)


def unfilter_rows(filtered: np.ndarray, rows: int, row_bytes: int, bpp: int) -> np.ndarray:
    filtered = np.ascontiguousarray(filtered, dtype=np.uint8).reshape(-1)
    out = np.empty(rows * row_bytes, np.uint8)
    rc = lib().r2po_unfilter_rows(_u8p(filtered), None, _u8p(out), rows, row_bytes, bpp)
    if rc:
        raise ValueError("invalid filter type byte")
    return out


def adler32(data, value: int = 1) -> int:
    a = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    a = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    return lib().r2po_adler32(value, _u8p(a), a.size)


def crc32(data, value: int = 0) -> int:
    a = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    a = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    return lib().r2po_crc32(value, _u8p(a), a.size)


def adler32_combine(a1: int, a2: int, len2: int) -> int:
    return lib().r2po_adler32_combine(a1, a2, len2)


def crc32_combine(c1: int, c2: int, len2: int) -> int:
    return lib().r2po_crc32_combine(c1, c2, len2)


@dataclass
class Encoded:
    png: bytes
    filtered: np.ndarray


MODE_READY, MODE_SWAP16, MODE_DICOM, MODE_DICOM8 = 0, 1, 2, 3


def encode_png(pixels: np.ndarray, width: int, height: int, channels: int, in_bits: int,
               mode: int, params: DicomParams | None = None, strip_height: int = 128,
               level: int = 0, threads: int = 1, want_filtered: bool = True) -> Encoded:
    """CPU baseline encode (window + filter + chunked zlib + PNG framing)."""
    pixels = np.ascontiguousarray(pixels)
    out_bits = in_bits if mode < 2 else (params.bits_allocated if mode == 2 else 8)
    fsize = lib().r2po_filtered_size(width, height, channels, out_bits)
    cap = fsize + fsize // 8 + 65536 + 12 * (fsize // 32768 + 4)
    out = np.empty(cap, np.uint8)
    filt = np.empty(fsize, np.uint8) if want_filtered else None
    n = lib().r2po_encode_png(_u8p(pixels.view(np.uint8)), width, height, channels, in_bits,
                              mode, C.byref(params) if params is not None else None,
                              strip_height, level, threads, _u8p(out), cap,
                              _u8p(filt) if filt is not None else None)
    if n == 0:
        raise RuntimeError("oracle encode failed")
    return Encoded(out[:n].tobytes(), filt)


def zlib_version() -> str:
    return lib().r2po_zlib_version().decode()
