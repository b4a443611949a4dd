"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Bit-exact for every integer/byte result; DEFLATE parity is defined on the
inflated bytes (stock zlib), never on the bitstream.
"""
import io
import sys
import zlib

import numpy as np
import pytest

from pngutil import check_and_inflate

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev(torch, arr: np.ndarray):
    t = torch.from_numpy(np.ascontiguousarray(arr).view(np.uint8).reshape(-1).copy())
    return t.cuda()


def empty(torch, n):
    return torch.empty(max(int(n), 1), dtype=torch.uint8, device="cuda")


def oracle_ready(oracle, r2p, px, img):
    """PNG-ready rows the reference would filter for this image."""
    flat = np.ascontiguousarray(px).reshape(-1)
    if img.pix_mode == r2p.PIX_READY:
        return flat.view(np.uint8)
    if img.pix_mode == r2p.PIX_SWAP16:
        return flat.astype("<u2").byteswap().view(np.uint8)
    p = oracle.DicomParams(*[getattr(img.dicom, f) for f, _ in img.dicom._fields_])
    if img.pix_mode == r2p.PIX_DICOM:
        return oracle.dicom_window(flat, p)
    return oracle.window8(flat.astype("<u2"), p)


def gpu_filter(torch, r2p, px, img, band_rows=0, prior=None):
    rb = r2p.row_bytes(img)
    d_src = dev(torch, px)
    d_out = empty(torch, img.height * (rb + 1))
    d_last = empty(torch, rb)
    d_prior = dev(torch, prior) if prior is not None else None
    r2p.filter_rows(img, d_src.data_ptr(), img.height, d_out.data_ptr(),
                    d_prior.data_ptr() if d_prior is not None else 0, d_last.data_ptr(), band_rows)
    torch.cuda.synchronize()
    return d_out.cpu().numpy(), d_last.cpu().numpy()[:rb]


CASES = [
    # (name, width, height, channels, in_bits, pix_mode)
    ("gray8", 37, 19, 1, 8, 0), ("gray8_w1", 1, 9, 1, 8, 0), ("gray8_h1", 100, 1, 1, 8, 0),
    ("ga8", 33, 7, 2, 8, 0), ("rgb8", 51, 13, 3, 8, 0), ("rgba8", 29, 11, 4, 8, 0),
    ("gray16be", 45, 9, 1, 16, 0), ("rgb16be", 23, 8, 3, 16, 0), ("rgba16be", 17, 6, 4, 16, 0),
    ("gray16swap", 64, 10, 1, 16, 1), ("rgb16swap", 21, 5, 3, 16, 1),
    ("rgb8_wide", 6000, 3, 3, 8, 0), ("gray16_wide", 8192, 3, 1, 16, 1),
    # rows over the 32 KiB that k_filter stages in LDS: the row-tiled kernel
    ("rgb16_huge", 17043, 7, 3, 16, 0), ("gray8_huge", 40000, 5, 1, 8, 0), ("rgb8_huge", 11000, 4, 3, 8, 0),
    ("rgba16_huge_swap", 9000, 3, 4, 16, 1), ("gray8_tile_edge", 16384 * 3, 2, 1, 8, 0),
]


@pytest.mark.parametrize("name,w,h,ch,bits,mode", CASES)
def test_filter_ready_modes(torch_dev, r2p, oracle, name, w, h, ch, bits, mode):
    log(f"case filter {name} {w}x{h}x{ch} bits={bits} mode={mode}")
    rng = np.random.default_rng(hash(name) & 0xFFFF)
    dt = np.uint8 if bits == 8 else "<u2"
    hi = 256 if bits == 8 else 65536
    px = rng.integers(0, hi, (h, w * ch)).astype(dt)
    px[:, : (w * ch) // 3] = px[0, 0]                 # flat run: filter ties
    img = r2p.Image.make(w, h, ch, bits, mode)
    out, last = gpu_filter(torch_dev, r2p, px, img)
    ready = oracle_ready(oracle, r2p, px, img)
    rb = r2p.row_bytes(img)
    bpp = ch * bits // 8
    exp = oracle.filter_rows(ready, h, rb, bpp)
    assert np.array_equal(out, exp)
    assert np.array_equal(last, ready[(h - 1) * rb:])


# rows up to 6 KiB: one row per wave; 6-32 KiB: rows spread over waves in
# persistent workgroups (band heights 1, 2, 4, 8); wider: tiled
@pytest.mark.parametrize("band_rows,w", [(1, 300), (3, 300), (16, 300), (0, 12000), (1, 3000), (3, 2500),
                                         (16, 4000), (0, 2200), (2, 10000)])
def test_filter_bands_and_prior(torch_dev, r2p, oracle, band_rows, w):
    log(f"case filter bands={band_rows} w={w}")
    rng = np.random.default_rng(band_rows)
    h = 41
    px = (np.cumsum(rng.integers(-4, 5, (h, w * 3)), axis=1) % 256).astype(np.uint8)
    prior = rng.integers(0, 256, w * 3, dtype=np.uint8)
    img = r2p.Image.make(w, h, 3, 8, 0)
    out, _ = gpu_filter(torch_dev, r2p, px, img, band_rows=band_rows, prior=prior)
    exp = oracle.filter_rows(px.reshape(-1), h, w * 3, 3, prior=prior)
    assert np.array_equal(out, exp)


def _dicom(r2p, *a):
    return r2p.DicomParams.make(*a)


DICOM_CASES = [
    ("ct_q15", 16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0, 3),
    ("ct_float16", 16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0, 2),
    ("signed_float16", 16, 12, 11, 1, 1, 2.0, -1024.0, 1, -600.0, 1500.0, 2),
    ("leftaligned_float16", 16, 12, 15, 0, 0, 1.0, 0.0, 1, 1000.5, 301.25, 2),
    ("narrow_window16", 16, 12, 11, 0, 0, 1.0, 0.0, 1, 100.0, 1.0, 2),
    ("nowindow16", 16, 14, 13, 0, 0, 1.0, 0.0, 0, 0.0, 0.0, 2),
    ("lut8_narrow", 16, 12, 11, 0, 0, 1.0, 0.0, 1, 40.0, 1.0, 3),
    ("lut8_shift", 16, 12, 15, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0, 3),
    ("q15_signed", 16, 12, 11, 1, 1, 1.0, -1024.0, 1, 40.0, 400.0, 3),
    ("q15_nowindow", 16, 12, 11, 0, 0, 1.0, 0.0, 0, 0.0, 0.0, 3),
    ("lut8_16bit", 16, 16, 15, 1, 1, 1.0, 0.0, 1, 0.0, 0.9, 3),
    ("dicom8_float", 8, 8, 7, 1, 1, 1.5, -3.0, 1, 0.0, 100.0, 2),
    # 16 bits stored: the 128 KiB window table is read from global memory
    ("bs16_float16", 16, 16, 15, 0, 1, 1.0, -1024.0, 1, 1000.0, 30000.0, 2),
    ("bs16_signed_float16", 16, 16, 15, 1, 1, 2.0, -100.0, 1, -300.0, 4000.0, 2),
    ("bs16_nowindow16", 16, 16, 15, 0, 0, 1.0, 0.0, 0, 0.0, 0.0, 2),
    ("bs16_signed_nowindow16", 16, 16, 15, 1, 1, 0.5, 10.0, 0, 0.0, 0.0, 2),
    ("bs16_narrow_window16", 16, 16, 15, 1, 0, 1.0, 0.0, 1, -5.0, 1.0, 2),
]


@pytest.mark.parametrize("wide", [False, True], ids=["rows512", "rows8k"])
@pytest.mark.parametrize("case", DICOM_CASES, ids=[c[0] for c in DICOM_CASES])
def test_window_exhaustive(torch_dev, r2p, oracle, case, wide):
    """Every stored code (plus garbage high bits) through the fused kernel,
    in rows of 512 samples and in 8 KiB rows (the kernel that spreads a row
    over several waves)."""
    name, ba, bs, hb, rep, ar, sl, ic, aw, wc, ww, mode = case
    log(f"case window {name} wide={wide}")
    p = _dicom(r2p, ba, bs, hb, rep, ar, sl, ic, aw, wc, ww)
    if ba == 8:
        codes = np.arange(256, dtype=np.uint8)
        w = 256
    else:
        n = 1 << 16 if bs == 16 or hb == 15 else 1 << bs
        codes = np.arange(n, dtype=np.uint32)
        if hb < 15 and bs < 16:
            codes = codes | (np.arange(n, dtype=np.uint32) % 7 << 13) * (bs <= 12)  # garbage high bits
        codes = codes.astype("<u2")
        w = 512
    if wide:
        w = 8192 if mode == 3 or ba == 8 else 4096      # 8 KiB PNG-ready rows
    h = (codes.size + w - 1) // w
    px = np.resize(codes, h * w).reshape(h, w)
    img = r2p.Image.make(w, h, 1, ba, mode, p)
    out, _ = gpu_filter(torch_dev, r2p, px, img)
    ready = oracle_ready(oracle, r2p, px, img)
    rb = r2p.row_bytes(img)
    bpp = r2p.out_bits(img) // 8
    got_ready = oracle.unfilter_rows(out, h, rb, bpp)
    assert np.array_equal(got_ready, ready)
    assert np.array_equal(out, oracle.filter_rows(ready, h, rb, bpp))


@pytest.mark.parametrize("n", [0, 1, 3, 15, 16, 17, 4095, 65535, 65536, 65537, 131071, 131072, 131073,
                               1 << 20, (1 << 22) + 7, (40 << 20) + 3, (160 << 20) + 11])
def test_checksums_vs_zlib(torch_dev, r2p, n):
    log(f"case checksum n={n}")
    rng = np.random.default_rng(n)
    data = rng.integers(0, 256, n, dtype=np.uint8)
    if n > 100:
        data[: n // 2] = 255                         # exercise the deferred-modulo bound
    d = dev(torch_dev, data) if n else empty(torch_dev, 1)
    work = empty(torch_dev, r2p.checksum_work_bytes(n))
    res = torch_dev.zeros(2, dtype=torch_dev.int64, device="cuda")
    r2p.adler32_device(d.data_ptr(), n, work.data_ptr(), res.data_ptr())
    r2p.crc32_device(d.data_ptr(), n, work.data_ptr(), res.data_ptr() + 8)
    torch_dev.cuda.synchronize()
    v = res.cpu().numpy().view(np.uint32)
    assert v[0] == zlib.adler32(data.tobytes())
    assert v[2] == zlib.crc32(data.tobytes())


def encode_device(torch, r2p, px, img, opt=None):
    opt = opt or r2p.options()
    fb = r2p.filtered_bytes(img)
    enc = r2p.Encoder(fb, opt, timing=True)
    d_src = dev(torch, px)
    cap = r2p.png_bound(img, opt)
    d_png = empty(torch, cap)
    d_filt = empty(torch, fb)
    enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap, d_filt.data_ptr())
    n, adler, nseg = enc.sync()
    png = d_png[:n].cpu().numpy().tobytes()
    filt = d_filt.cpu().numpy()
    t = enc.timings()
    enc.close()
    return png, filt, adler, nseg, t


def _check_png(oracle, r2p, png, px, img, filt_gpu):
    from PIL import Image as PILImage
    hdr, z, raw = check_and_inflate(png)
    ready = oracle_ready(oracle, r2p, px, img)
    rb = r2p.row_bytes(img)
    bpp = rb // img.width
    exp = oracle.filter_rows(ready, img.height, rb, bpp)
    assert raw == exp.tobytes(), "inflated IDAT differs from the oracle's filtered bytes"
    assert np.array_equal(filt_gpu, exp)
    assert hdr["width"] == img.width and hdr["height"] == img.height
    assert hdr["depth"] == r2p.out_bits(img)
    ob = r2p.out_bits(img)
    if not (ob == 16 and img.channels > 1):          # (PIL reads 16-bit colour as 8-bit)
        im = np.array(PILImage.open(io.BytesIO(png)))
        expect = ready.view(">u2") if ob == 16 else ready
        assert np.array_equal(im.reshape(-1).astype(np.int64), expect.astype(np.int64))
    return z


@pytest.mark.parametrize("lz77", [0, 1, 2])
@pytest.mark.parametrize("kind", ["ct_slice", "radiograph_small", "camera_small"])
def test_encode_png_workloads(torch_dev, r2p, synth, oracle, kind, lz77):
    log(f"case encode {kind} lz77={lz77}")
    if kind == "ct_slice":
        s = synth.ct_slice(0)
    elif kind == "radiograph_small":
        s = synth.radiograph(1, size=640)
    else:
        s = synth.camera_rgb(2, 900, 700)
    opt = r2p.options(lz77=lz77)
    png, filt, adler, nseg, t = encode_device(torch_dev, r2p, s.pixels, s.img, opt)
    z = _check_png(oracle, r2p, png, s.pixels, s.img, filt)
    assert adler == zlib.adler32(filt.tobytes())


@pytest.mark.parametrize("w,h,ch,bits,mode,dicom", [
    (17043, 9, 3, 16, 0, None),                                               # RGB16, 102 KB rows
    (40000, 6, 1, 8, 0, None),                                                # gray8, 40 KB rows
    (20000, 5, 1, 16, 2, (16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)),   # float window, 40 KB rows
    (36000, 4, 1, 16, 3, (16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)),   # Q15 16 -> 8, 36 KB rows
    (18000, 3, 1, 16, 2, (16, 16, 15, 1, 1, 1.0, 0.0, 1, 0.0, 3000.0)),       # 16 bits stored, global table
])
def test_encode_huge_rows(torch_dev, r2p, oracle, w, h, ch, bits, mode, dicom):
    """Rows wider than k_filter's LDS rows through the whole encoder."""
    log(f"case encode huge rows {w}x{h}x{ch} bits={bits} mode={mode}")
    rng = np.random.default_rng(w + h)
    hi = 256 if bits == 8 else (4096 if dicom and dicom[1] == 12 else 65536)
    px = (np.cumsum(rng.integers(-3, 4, (h, w * ch)), axis=1) % hi).astype(np.uint8 if bits == 8 else "<u2")
    px[:, : (w * ch) // 4] = rng.integers(0, hi, (h, (w * ch) // 4))
    p = r2p.DicomParams.make(*dicom) if dicom else None
    img = r2p.Image.make(w, h, ch, bits, mode, p) if p else r2p.Image.make(w, h, ch, bits, mode)
    for lz77 in (0, 1, 2):
        png, filt, adler, nseg, t = encode_device(torch_dev, r2p, px, img, r2p.options(lz77=lz77))
        _check_png(oracle, r2p, png, px, img, filt)
        assert adler == zlib.adler32(filt.tobytes())


def test_lz77_finds_repeats(torch_dev, r2p, oracle):
    """Noise-free tiled texture, flat areas and gradients: LZ77 must shrink them."""
    log("case lz77 repeats")
    rng = np.random.default_rng(6)
    w, h = 1000, 300
    tile = rng.integers(0, 256, (32, 32, 3), dtype=np.uint8)
    px = np.tile(tile, (h // 32 + 1, w // 32 + 1, 1))[:h, :w].copy()
    px[200:, :, :] = 77
    img = r2p.Image.make(w, h, 3, 8, 0)
    png0, filt, _, _, _ = encode_device(torch_dev, r2p, px, img, r2p.options(lz77=0))
    png1, filt1, _, _, _ = encode_device(torch_dev, r2p, px, img, r2p.options(lz77=1))
    z1 = _check_png(oracle, r2p, png1, px, img, filt1)
    assert len(png1) < 0.35 * len(png0), (len(png1), len(png0))
    # the same through the dynamic-Huffman blocks (long matches in the tokens)
    png2, filt2, _, _, _ = encode_device(torch_dev, r2p, px, img, r2p.options(lz77=1, huffman=1))
    _check_png(oracle, r2p, png2, px, img, filt2)
    assert len(png2) < len(png1)
    # the denser search (every 2nd position) finds at least as much here
    png3, filt3, _, _, _ = encode_device(torch_dev, r2p, px, img, r2p.options(lz77=2))
    _check_png(oracle, r2p, png3, px, img, filt3)
    assert len(png3) <= len(png1) * 1.01


@pytest.mark.parametrize("seg,chunk", [(4096, 1000), (65536, 65536), (1 << 20, 8192)])
def test_encode_segment_and_chunk_sizes(torch_dev, r2p, oracle, seg, chunk):
    log(f"case encode seg={seg} chunk={chunk}")
    rng = np.random.default_rng(seg)
    w, h = 777, 233
    px = (np.cumsum(rng.integers(-2, 3, (h, w * 3)), axis=1) % 256).astype(np.uint8)
    px[50:120, :] = 17
    img = r2p.Image.make(w, h, 3, 8, 0)
    opt = r2p.options(segment_bytes=seg, idat_chunk_bytes=chunk)
    png, filt, adler, nseg, t = encode_device(torch_dev, r2p, px, img, opt)
    _check_png(oracle, r2p, png, px, img, filt)


@pytest.mark.parametrize("seg,chunk,w,h", [
    (4096, 333, 1500, 700),        # 257 segments: self-scanned assembly
    (4096, 777, 3000, 2000),       # 1466 segments: assembly after its own scan launch
    (8192, 65536, 1500, 700),      # two slices per segment, full-size chunks
])
def test_assembly_boundaries(torch_dev, r2p, oracle, seg, chunk, w, h):
    """Incompressible rows with small segments and odd IDAT chunk sizes: slice,
    segment and chunk boundaries every few hundred bytes, so most waves of the
    assembly gather 16-byte groups across them."""
    log(f"case assembly boundaries seg={seg} chunk={chunk} {w}x{h}")
    rng = np.random.default_rng(seg + chunk)
    px = rng.integers(0, 256, (h, w), dtype=np.uint8)
    px[h // 2: h // 2 + 40, :] = 9                 # a compressible band between random ones
    img = r2p.Image.make(w, h, 1, 8, 0)
    for lz77 in (0, 1):
        opt = r2p.options(segment_bytes=seg, slice_bytes=4096, idat_chunk_bytes=chunk, lz77=lz77)
        png, filt, adler, nseg, t = encode_device(torch_dev, r2p, px, img, opt)
        _check_png(oracle, r2p, png, px, img, filt)
        assert adler == zlib.adler32(filt.tobytes())


def test_adler_across_chunks(torch_dev, r2p, oracle):
    """Chunk sizes that end the zlib stream's first IDAT chunk 1..5 bytes
    before the stream's end, so that the Adler-32 trailer is split across two
    chunks (at 1 byte the last chunk holds a single byte)."""
    log("case adler across chunks")
    rng = np.random.default_rng(11)
    w, h = 300, 200
    px = (np.cumsum(rng.integers(-2, 3, (h, w)), axis=1) % 256).astype(np.uint8)
    img = r2p.Image.make(w, h, 1, 8, 0)
    png, filt, _, _, _ = encode_device(torch_dev, r2p, px, img)
    z = len(check_and_inflate(png)[1])                  # zlib stream bytes
    assert z <= 65536 + 5
    for chunk in [z - 1, z - 2, z - 3, z - 4, z - 5]:
        opt = r2p.options(idat_chunk_bytes=chunk)
        png2, filt2, adler, _, _ = encode_device(torch_dev, r2p, px, img, opt)
        _check_png(oracle, r2p, png2, px, img, filt2)
        assert adler == zlib.adler32(filt2.tobytes())


@pytest.mark.parametrize("seg,slc,w,h", [
    (65536, 4096, 2340, 7),        # 16387 bytes: four full 4 KiB slices and one of 3 bytes
    (8192, 4096, 2340, 7),         # three segments, the last one a 3-byte slice
    (65536, 16384, 1500, 300),     # segments of four slices, ragged last segment
    (65536, 65536, 1500, 300),     # one slice per segment
    (65536, 4096, 4096, 2047),     # 2048 slices: the most the assembly scans itself
    (65536, 4096, 4000, 2200),     # 2149 slices: scanned by its own launch
    (4096, 4096, 3000, 2000),      # 1466 one-slice segments: over the self-scan's 1024
])
def test_encode_slices(torch_dev, r2p, oracle, seg, slc, w, h):
    """Slices are a work split only: any slice size inflates to the same bytes,
    including segments whose last slice holds a few bytes."""
    log(f"case encode seg={seg} slice={slc} {w}x{h}")
    rng = np.random.default_rng(w + h)
    px = (np.cumsum(rng.integers(-3, 4, (h, w)), axis=1) % 256).astype(np.uint8)
    px[h // 3: h // 2, :] = 200
    img = r2p.Image.make(w, h, 1, 8, 0)
    for lz77 in (0, 1, 2):
        opt = r2p.options(segment_bytes=seg, slice_bytes=slc, lz77=lz77)
        assert (opt.segment_bytes, opt.slice_bytes) == (seg, slc)
        png, filt, adler, nseg, t = encode_device(torch_dev, r2p, px, img, opt)
        _check_png(oracle, r2p, png, px, img, filt)
        assert adler == zlib.adler32(filt.tobytes())


def test_encode_tiny_and_host_path(torch_dev, r2p, oracle):
    log("case encode tiny/host")
    for (w, h, ch, bits) in [(1, 1, 1, 8), (2, 1, 3, 8), (1, 3, 1, 16), (5, 4, 4, 16)]:
        rng = np.random.default_rng(w * 10 + h)
        px = rng.integers(0, 256 if bits == 8 else 65536, (h, w * ch)).astype(np.uint8 if bits == 8 else ">u2")
        img = r2p.Image.make(w, h, ch, bits, 0)
        enc = r2p.Encoder(r2p.filtered_bytes(img))
        png = enc.encode_host(img, px)
        enc.close()
        hdr, z, raw = check_and_inflate(png)
        ready = px.reshape(-1).view(np.uint8)
        rb = r2p.row_bytes(img)
        assert raw == oracle.filter_rows(ready, h, rb, rb // w).tobytes()


def test_strip_deflate_concatenation(torch_dev, r2p):
    log("case strip deflate")
    rng = np.random.default_rng(11)
    data = (np.cumsum(rng.integers(-1, 2, 600000)) % 256).astype(np.uint8)
    data[100000:200000] = 0
    sd = r2p.StripDeflate(200000, data.size, lz77=1, segment_bytes=65536)
    d = dev(torch_dev, data)
    cuts = [0, 1, 70000, 200000, 200001, 400000, 600000]
    for i in range(len(cuts) - 1):
        a, b = cuts[i], cuts[i + 1]
        sd.strip(d.data_ptr() + a, b - a, is_last=(i == len(cuts) - 2), bpp=1)
        sd.sync()
    raw = sd.output()
    st = sd.lz_stats()
    assert zlib.decompress(raw, -15) == data.tobytes()
    assert sd.adler() == zlib.adler32(data.tobytes())
    assert st.matches > 0 and st.matched_bytes > 100000
    assert st.max_len > 16                        # runs merged across parse lanes (<= 258)
    assert st.max_len <= 258
    # a stream whose strips are all non-final is a valid sync-flushed prefix
    sd.reset()
    sd.strip(d.data_ptr(), 1000, is_last=False)
    sd.sync()
    part = sd.output()
    do = zlib.decompressobj(-15)
    assert do.decompress(part) == data[:1000].tobytes()
    sd.close()


def test_idat_chunk_device(torch_dev, r2p):
    log("case idat chunk")
    rng = np.random.default_rng(12)
    payload = rng.integers(0, 256, 70000, dtype=np.uint8)
    d = dev(torch_dev, payload)
    cap = payload.size + 64
    out = empty(torch_dev, cap)
    work = empty(torch_dev, r2p.checksum_work_bytes(cap))
    for first, last in [(1, 0), (0, 0), (0, 1), (1, 1)]:
        n = r2p.idat_chunk(d.data_ptr(), payload.size, first, last, 0x11E60398, 6, out.data_ptr(),
                           cap, work.data_ptr())
        torch_dev.cuda.synchronize()
        b = out[:n].cpu().numpy().tobytes()
        data_len = int.from_bytes(b[:4], "big")
        assert b[4:8] == b"IDAT" and data_len == n - 12
        body = b[8:8 + data_len]
        exp = (b"\x78\x9c" if first else b"") + payload.tobytes() + (b"\x11\xe6\x03\x98" if last else b"")
        assert body == exp
        assert int.from_bytes(b[-4:], "big") == zlib.crc32(b[4:8] + body)
