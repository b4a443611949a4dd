"""Pins the CPU oracle against known answers, closed forms, zlib, PIL and the
reference's own zlib-framing sources (oracle/_ref, when built).  CPU only."""
import io
import os
import tempfile
import zlib

import numpy as np
import pytest

from pngutil import check_and_inflate


def test_q15_known_answer_ct_slice(oracle):
    # CT slice: slope 1, intercept -1024, WC 40, WW 400 -> A = 20889, B = -18048614
    p = oracle.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)
    assert oracle.q15_coeffs(p) == (20889, -18048614)
    codes = np.arange(4096, dtype="<u2")
    out = oracle.window8(codes, p)
    # closed form of the linear window in exact arithmetic: (HU - lo) * 255 / WW
    hu = codes.astype(np.float64) - 1024.0
    exact = (hu - (40.0 - 200.0)) * 255.0 / 400.0
    assert np.all(np.abs(out.astype(np.float64) - np.clip(np.floor(exact), 0, 255)) <= 1)
    assert np.all(out[codes <= 865] == 0) and np.all(out[codes >= 1265] == 255)
    assert np.all(np.diff(out.astype(int)) >= 0)


def test_q15_rejections(oracle):
    p = oracle.DicomParams.make(16, 12, 15, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)  # shift != 0
    assert oracle.q15_coeffs(p) is None
    p = oracle.DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 1, 40.0, 1.0)  # A*32768 overflows
    assert oracle.q15_coeffs(p) is None
    # rejected parameters fall back to the LUT, whose closed form we check
    codes = np.arange(4096, dtype="<u2")
    out = oracle.window8(codes, p)
    lo, hi = 40.0 - 0.5, 40.0 + 0.5
    exact = np.floor((codes - lo) / (hi - lo) * 255.0 + 0.5)
    assert np.array_equal(out, np.clip(exact, 0, 255).astype(np.uint8))


def test_q15_no_window_full_range(oracle):
    p = oracle.DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 0)
    a, b = oracle.q15_coeffs(p)
    codes = np.arange(4096, dtype="<u2")
    out = oracle.window8(codes, p).astype(int)
    exact = np.floor(codes * 255.0 / 4095.0)
    # A = trunc(255/4095 * 32768) = 2040 truncates, so the top code maps to 254
    # (4095 * 2040 >> 15), exactly as the reference's Q15 loop computes it.
    assert (a, b) == (2040, 0)
    assert out[0] == 0 and out[-1] == 254
    assert np.all(np.abs(out - exact) <= 1)


def test_lut8_signed(oracle):
    p = oracle.DicomParams.make(16, 12, 11, 1, 0, 1.0, 0.0, 0)
    lut = oracle.lut8(p)
    # signed 12-bit: code 2048 is -2048 -> 0; code 2047 is +2047 -> 255
    assert lut[2048] == 0 and lut[2047] == 255
    vals = np.where(np.arange(4096) >= 2048, np.arange(4096) - 4096, np.arange(4096))
    exact = np.floor((vals + 2048) / 4096.0 * 255.0 + 0.5)
    assert np.array_equal(lut, np.clip(exact, 0, 255).astype(np.uint8))


def _ps33_exact(v, c, w, ymax):
    v = np.asarray(v, np.float64)
    wm1 = w - 1.0
    y = ((v - c + 0.5) / wm1 + 0.5) * ymax
    y = np.where(v <= c - 0.5 - wm1 / 2, 0, np.where(v > c - 0.5 + wm1 / 2, ymax, y))
    return np.clip(y, 0, ymax)


@pytest.mark.parametrize("c,w,slope,icpt,signed", [
    (40.0, 400.0, 1.0, -1024.0, 0), (2048.0, 4096.0, 1.0, 0.0, 0), (1000.5, 301.25, 1.0, 0.0, 0),
    (-600.0, 1500.0, 2.0, -1024.0, 1), (100.0, 1.0, 1.0, 0.0, 0), (0.0, 0.5, 1.0, 0.0, 1)])
def test_float_window_closed_form(oracle, c, w, slope, icpt, signed):
    p = oracle.DicomParams.make(16, 12, 11, signed, 1, slope, icpt, 1, c, w)
    codes = np.arange(4096, dtype="<u2")
    be = oracle.dicom_window(codes, p).view(">u2").astype(np.int64)
    stored = codes.astype(np.int64)
    if signed:
        stored = np.where(stored >= 2048, stored - 4096, stored)
    v = stored * slope + icpt
    if w - 1.0 <= 0:
        exp = np.where(v >= c, 65535, 0)
        assert np.array_equal(be, exp)
        return
    exact = _ps33_exact(v, c, w, 65535.0)
    assert np.all(np.abs(be - np.floor(exact)) <= 1)
    assert np.mean(be == np.floor(exact)) > 0.97
    assert np.all(be[v <= c - 0.5 - (w - 1) / 2] == 0)
    assert np.all(be[v > c - 0.5 + (w - 1) / 2] == 65535)


@pytest.mark.parametrize("signed,rescale,slope,icpt,window,c,w", [
    (0, 0, 1.0, 0.0, 0, 0.0, 0.0), (1, 0, 1.0, 0.0, 0, 0.0, 0.0), (1, 1, 2.0, -100.0, 0, 0.0, 0.0),
    (0, 1, 1.0, -1024.0, 1, 1000.0, 30000.0), (1, 1, 2.0, -100.0, 1, -300.0, 4000.0),
    (1, 0, 1.0, 0.0, 1, -5.0, 1.0)])
def test_float_window_16_bits_stored(oracle, signed, rescale, slope, icpt, window, c, w):
    """All 65536 codes of 16-bit stored samples (no shift, no mask) against the
    PS3.3 closed form: identity without window/rescale (negative values clamp
    to 0), the window's special cases exactly, the formula to +-1."""
    p = oracle.DicomParams.make(16, 16, 15, signed, rescale, slope, icpt, window, c, w)
    codes = np.arange(65536, dtype=np.uint32).astype("<u2")
    be = oracle.dicom_window(codes, p).view(">u2").astype(np.int64)
    stored = codes.astype(np.int64)
    if signed:
        stored = np.where(stored >= 32768, stored - 65536, stored)
    v = stored * slope + icpt if rescale else stored.astype(np.float64)
    if not window:
        assert np.array_equal(be, np.clip(np.floor(v), 0, 65535).astype(np.int64))
        return
    if w - 1.0 <= 0:
        assert np.array_equal(be, np.where(v >= c, 65535, 0))
        return
    exact = _ps33_exact(v, c, w, 65535.0)
    assert np.all(np.abs(be - np.floor(exact)) <= 1)
    assert np.mean(be == np.floor(exact)) > 0.97
    assert np.all(be[v <= c - 0.5 - (w - 1) / 2] == 0)
    assert np.all(be[v > c - 0.5 + (w - 1) / 2] == 65535)


def test_float_window_shift_mask_and_8bit(oracle):
    # left-aligned 12-bit data (high bit 15): stored value = raw >> 4; garbage low bits ignored
    p = oracle.DicomParams.make(16, 12, 15, 0, 0, 1.0, 0.0, 0)
    raw = (np.arange(4096, dtype=np.uint32) << 4 | 0x5).astype("<u2")
    out = oracle.dicom_window(raw, p).view(">u2")
    assert np.array_equal(out, np.arange(4096))     # no window: identity after shift+mask
    # high garbage bits above bits_stored are masked
    p = oracle.DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 0)
    raw = (np.arange(4096, dtype=np.uint32) | 0xF000).astype("<u2")
    assert np.array_equal(oracle.dicom_window(raw, p).view(">u2"), np.arange(4096))
    # 8-bit allocated, signed, window
    p = oracle.DicomParams.make(8, 8, 7, 1, 0, 1.0, 0.0, 1, 0.0, 100.0)
    raw = np.arange(256, dtype=np.uint8)
    out = oracle.dicom_window(raw, p).astype(np.int64)
    v = raw.astype(np.int8).astype(np.float64)
    assert np.all(np.abs(out - np.floor(_ps33_exact(v, 0.0, 100.0, 255.0))) <= 1)


def _filters_bruteforce(cur, prev, bpp):
    """Filter definitions of the PNG specification (section 9.2), byte by byte."""
    n = len(cur)
    res = []
    for f in range(5):
        row = []
        for i in range(n):
            x, b = cur[i], prev[i]
            a = cur[i - bpp] if i >= bpp else 0
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0: pred = 0
            elif f == 1: pred = a
            elif f == 2: pred = b
            elif f == 3: pred = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            row.append((x - pred) % 256)
        res.append(row)
    return res


@pytest.mark.parametrize("bpp,width,rows", [(1, 7, 5), (2, 5, 4), (3, 6, 3), (4, 3, 3), (6, 2, 4), (8, 3, 2)])
def test_filter_bruteforce_tiny(oracle, bpp, width, rows):
    rng = np.random.default_rng(bpp * 100 + width)
    rb = width * bpp
    for trial in range(20):
        img = rng.integers(0, 256, (rows, rb), dtype=np.uint8)
        if trial % 3 == 0:
            img[:, : rb // 2] = img[0, 0]          # flat area: ties between filters
        out = oracle.filter_rows(img, rows, rb, bpp).reshape(rows, rb + 1)
        for r in range(rows):
            prev = img[r - 1].tolist() if r else [0] * rb
            cands = _filters_bruteforce(img[r].tolist(), prev, bpp)
            cost = [sum(min(v, 256 - v) for v in c) for c in cands]
            best = int(np.argmin(cost))            # argmin returns the lowest index on ties
            assert out[r, 0] == best
            assert out[r, 1:].tolist() == cands[best]


def test_filter_roundtrip_and_pil(oracle):
    from PIL import Image as PILImage
    rng = np.random.default_rng(7)
    img = (np.cumsum(rng.integers(-3, 4, (37, 53, 3)), axis=1) % 256).astype(np.uint8)
    filt = oracle.filter_rows(img.reshape(-1), 37, 159, 3)
    assert np.array_equal(oracle.unfilter_rows(filt, 37, 159, 3), img.reshape(-1))
    e = oracle.encode_png(img, 53, 37, 3, 8, oracle.MODE_READY, strip_height=8, level=1, threads=2)
    hdr, z, raw = check_and_inflate(e.png)
    assert raw == filt.tobytes()
    assert np.array_equal(np.array(PILImage.open(io.BytesIO(e.png))), img)


def test_checksums_known_answers(oracle):
    assert oracle.adler32(b"Wikipedia") == 0x11E60398
    assert oracle.crc32(b"123456789") == 0xCBF43926
    assert oracle.crc32(b"IEND") == 0xAE426082
    assert oracle.adler32(b"") == 1 and oracle.crc32(b"") == 0
    rng = np.random.default_rng(1)
    for n in [1, 5551, 5552, 5553, 100000]:
        d = rng.integers(0, 256, n, dtype=np.uint8)
        assert oracle.adler32(d) == zlib.adler32(d.tobytes())
        assert oracle.crc32(d) == zlib.crc32(d.tobytes())
    d = np.full(200000, 255, np.uint8)                 # worst case for the deferred modulo
    assert oracle.adler32(d) == zlib.adler32(d.tobytes())


def test_checksum_combine(oracle):
    rng = np.random.default_rng(2)
    d = rng.integers(0, 256, 300001, dtype=np.uint8)
    for cut in [0, 1, 65536, 123457, 300001]:
        a, b = d[:cut], d[cut:]
        assert oracle.adler32_combine(oracle.adler32(a), oracle.adler32(b), len(b)) == zlib.adler32(d.tobytes())
        assert oracle.crc32_combine(oracle.crc32(a), oracle.crc32(b), len(b)) == zlib.crc32(d.tobytes())


def test_zlib_flg(oracle):
    for lvl in range(10):
        flg = oracle.lib().r2po_zlib_flg(lvl)
        assert (0x78 * 256 + flg) % 31 == 0
    assert [oracle.lib().r2po_zlib_flg(l) for l in (0, 1, 3, 6, 9)] == [0xDA, 0x01, 0x5E, 0x9C, 0xDA]


@pytest.mark.parametrize("mode", ["ct8", "ct16", "rgb"])
def test_cpu_baseline_png(oracle, mode):
    from PIL import Image as PILImage
    rng = np.random.default_rng(3)
    p = oracle.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)
    if mode == "rgb":
        px = rng.integers(0, 256, (45, 70, 3), dtype=np.uint8)
        e = oracle.encode_png(px, 70, 45, 3, 8, oracle.MODE_READY, strip_height=16, level=3, threads=3)
        expect = px
    else:
        px = rng.integers(0, 4096, (45, 70)).astype("<u2")
        m = oracle.MODE_DICOM8 if mode == "ct8" else oracle.MODE_DICOM
        e = oracle.encode_png(px, 70, 45, 1, 16, m, p, strip_height=16, level=0, threads=3)
        expect = oracle.window8(px.reshape(-1), p).reshape(45, 70) if mode == "ct8" else \
            oracle.dicom_window(px.reshape(-1), p).view(">u2").reshape(45, 70)
    hdr, z, raw = check_and_inflate(e.png)
    assert raw == e.filtered.tobytes()
    got = np.array(PILImage.open(io.BytesIO(e.png)))
    assert np.array_equal(got.astype(np.int64), expect.astype(np.int64))


# ---------------------------------------------------------------------------
# against the reference's own framing sources (oracle/_ref/libref.so)
# ---------------------------------------------------------------------------
def _ref_or_skip(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference tree at build time)")
    return R


def test_ref_zlib_header_trailer(oracle):
    import ctypes as C
    R = _ref_or_skip(oracle)
    for lvl in range(10):
        b = (C.c_uint8 * 2)()
        R.ref_zlib_header(lvl, b)
        assert bytes(b) == bytes([0x78, oracle.lib().r2po_zlib_flg(lvl)])
    t = (C.c_uint8 * 4)()
    R.ref_zlib_trailer(0x11E60398, t)
    assert bytes(t) == bytes([0x11, 0xE6, 0x03, 0x98])


def test_ref_deflate_strip_matches_oracle_encoder(oracle):
    """The oracle's chunked deflate equals deflate_strip byte for byte."""
    import ctypes as C
    R = _ref_or_skip(oracle)
    rng = np.random.default_rng(4)
    px = (np.cumsum(rng.integers(-2, 3, (300, 400)), axis=1) % 256).astype(np.uint8)
    for level, threads, sh in [(0, 6, 128), (1, 3, 64), (3, 6, 300)]:
        e = oracle.encode_png(px, 400, 300, 1, 8, oracle.MODE_READY, strip_height=sh, level=level,
                              threads=threads)
        hdr, z, raw = check_and_inflate(e.png)
        # re-deflate the same filtered strips through the reference
        filt = e.filtered
        rowb = 401
        parts, adlers, sizes = [], [], []
        nstrips = (300 + sh - 1) // sh
        for s in range(nstrips):
            a = s * sh * rowb
            b = min(300, (s + 1) * sh) * rowb
            buf = (C.c_uint8 * (2 * (b - a) + 4096))()
            ad, ov = C.c_uint32(), C.c_int()
            n = R.ref_deflate_strip(filt[a:b].ctypes.data_as(C.POINTER(C.c_uint8)), b - a, level,
                                    threads, int(s == nstrips - 1), buf, len(buf), C.byref(ad),
                                    C.byref(ov))
            assert ov.value == 0
            parts.append(bytes(buf[:n])); adlers.append(ad.value); sizes.append(b - a)
        A = (C.c_uint32 * len(adlers))(*adlers)
        S = (C.c_uint64 * len(sizes))(*sizes)
        run = R.ref_accum_adler(A, S, len(adlers))
        ref_stream = bytes([0x78, oracle.lib().r2po_zlib_flg(level)]) + b"".join(parts) + run.to_bytes(4, "big")
        assert ref_stream == z


def test_ref_png_writer_framing(oracle):
    R = _ref_or_skip(oracle)
    rng = np.random.default_rng(5)
    px = rng.integers(0, 256, (120, 333), dtype=np.uint8)
    e = oracle.encode_png(px, 333, 120, 1, 8, oracle.MODE_READY, strip_height=50, level=0, threads=2)
    hdr, z, raw = check_and_inflate(e.png)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "ref.png")
        zs = np.frombuffer(z, np.uint8)
        import ctypes as C
        assert R.ref_png_write(path.encode(), 333, 120, 8, 0, zs.ctypes.data_as(C.POINTER(C.c_uint8)), len(z))
        assert open(path, "rb").read() == e.png
