"""Parity at the benchmark workloads' full sizes, through the C-ABI: the
3072x3072 radiograph and the 6000x4000 RGB camera image against the oracle
byte for byte (filtered bytes and the inflated IDAT stream), the camera image
also as a TIFF file through encode_tiff_to_png, the full 1024-image mixed
batch against the plain form (every PNG) and the oracle (32 sampled images),
and the served form on the mixed request stream with 8192x8192 requests,
every PNG compared byte for byte with the plain form's."""
import os
import subprocess
import sys
import zlib

import numpy as np
import pytest

from conftest import REPO
from pngutil import check_and_inflate

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _params(oracle, img):
    return oracle.DicomParams(*[getattr(img.dicom, f) for f, _ in img.dicom._fields_])


def oracle_filtered(oracle, r2p, img, px):
    flat = np.ascontiguousarray(px).reshape(-1)
    if img.pix_mode == r2p.PIX_DICOM8:
        ready = oracle.window8(flat.view("<u2"), _params(oracle, img))
    elif img.pix_mode == r2p.PIX_DICOM:
        ready = oracle.dicom_window(flat.view("<u2") if img.in_bits == 16 else flat, _params(oracle, img))
    else:
        ready = flat.view(np.uint8)
    rb = r2p.row_bytes(img)
    return oracle.filter_rows(ready, img.height, rb, rb // img.width)


def encode_plain(torch, r2p, img, d_src, opt=None, want_filtered=False):
    opt = opt or r2p.options()
    fb = r2p.filtered_bytes(img)
    enc = r2p.Encoder(fb, opt)
    cap = r2p.png_bound(img, opt)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    d_filt = torch.empty(fb, dtype=torch.uint8, device="cuda") if want_filtered else None
    enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap, d_filt.data_ptr() if want_filtered else 0)
    n, adler, _ = enc.sync()
    png = d_png[:n].cpu().numpy().tobytes()
    enc.close()
    return png, (d_filt.cpu().numpy() if want_filtered else None), adler


@pytest.mark.timeout(300)
@pytest.mark.parametrize("kind", ["radiograph", "camera_rgb"])
def test_full_size_workload_vs_oracle(torch_dev, r2p, synth, oracle, kind):
    torch = torch_dev
    s = synth.workload(kind)
    log(f"case full-size {kind} {s.img.width}x{s.img.height}")
    d_src = torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda()
    exp = oracle_filtered(oracle, r2p, s.img, s.pixels)
    # automatic slices: the stream's last partial round of 1024 slices is
    # cut into half-size slices (1024 + 257 for the radiograph); explicit
    # 16 KiB slices keep every slice whole; dynamic-Huffman blocks over both
    for opt in (r2p.options(lz77=1), r2p.options(lz77=0), r2p.options(lz77=1, slice_bytes=16384),
                r2p.options(lz77=1, huffman=1)):
        png, filt, adler = encode_plain(torch, r2p, s.img, d_src, opt, want_filtered=True)
        assert np.array_equal(filt, exp), "filtered scanlines differ from the oracle"
        hdr, z, raw = check_and_inflate(png)
        assert raw == exp.tobytes(), "inflated IDAT differs from the oracle's filtered bytes"
        assert adler == zlib.adler32(raw)
        assert (hdr["width"], hdr["height"], hdr["depth"]) == (s.img.width, s.img.height, r2p.out_bits(s.img))
    # host -> host form: the same bytes
    enc = r2p.Encoder(r2p.filtered_bytes(s.img))
    assert enc.encode_host(s.img, s.pixels) == encode_plain(torch, r2p, s.img, d_src)[0]
    enc.close()


@pytest.mark.timeout(300)
def test_camera_tiff_file_full_size(torch_dev, r2p, synth, oracle, tmp_path):
    """The 6000x4000 RGB8 workload as an uncompressed little-endian TIFF through
    encode_tiff_to_png (GPU DEFLATE and the CPU-deflate form)."""
    exe = os.path.join(REPO, "tests", "cpp", "facade_check")
    s = synth.camera_rgb(2)
    log("case full-size camera TIFF 6000x4000")
    tif = str(tmp_path / "cam.tif")
    synth.write_tiff(tif, s.pixels, rows_per_strip=64)
    exp = oracle_filtered(oracle, r2p, s.img, s.pixels).tobytes()
    for gpu_deflate in (1, 0):
        out = str(tmp_path / f"cam{gpu_deflate}.png")
        r = subprocess.run([exe, "encode", "tiff", tif, out, str(gpu_deflate)], capture_output=True, text=True,
                           timeout=200)
        assert r.returncode == 0, r.stdout + r.stderr
        hdr, z, raw = check_and_inflate(open(out, "rb").read())
        assert raw == exp, f"gpu_deflate={gpu_deflate}: inflated bytes differ from the oracle"


@pytest.mark.timeout(600)
def test_full_mixed_batch_1024(torch_dev, r2p, synth, oracle):
    """The mixed batch of 1024 images (sides 256..4096, 50 % CT phantoms with
    8-bit output, 30 % radiographs with 16-bit output, 20 % RGB crops) as the
    benchmark generates it: every PNG of one graph replay equals the plain
    form's, and 32 sampled images inflate to the oracle's filtered bytes."""
    torch = torch_dev
    log("case full mixed batch 1024")
    items = synth.mixed_batch_device(torch, 1024, seed=3, max_side=4096)
    torch.cuda.synchronize()
    imgs = [im for im, _ in items]
    fsum = sum(r2p.filtered_bytes(im) for im in imgs)
    isum = sum(r2p.input_bytes(im) for im in imgs)
    b = r2p.Batch(len(imgs), isum, fsum, r2p.options())
    b.set(imgs, [t.data_ptr() for _, t in items])
    b.run(True)
    b.sync()
    b.run(True)                                        # a second replay: same bytes
    b.sync()
    opt = b.opt
    maxf = max(r2p.filtered_bytes(im) for im in imgs)
    enc = r2p.Encoder(maxf, opt)
    cap = max(r2p.png_bound(im, opt) for im in imgs)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    rng = np.random.default_rng(3)
    sample = set(rng.choice(len(imgs), size=32, replace=False).tolist())
    log(f"batch set: {fsum / 1e9:.2f} GB filtered; sampled {sorted(sample)[:8]}...")
    for i, (im, t) in enumerate(items):
        enc.encode_device(im, t.data_ptr(), d_png.data_ptr(), cap)
        n, _, _ = enc.sync()
        got = b.copy_result(i)
        assert got == d_png[:n].cpu().numpy().tobytes(), f"image {i} ({im.width}^2 mode {im.pix_mode}) differs"
        if i in sample:
            hdr, z, raw = check_and_inflate(got)
            px = t.cpu().numpy()
            assert raw == oracle_filtered(oracle, r2p, im, px).tobytes(), f"image {i} differs from the oracle"
        if i % 128 == 127:
            log(f"  {i + 1} images compared")
    enc.close()
    b.close()


@pytest.mark.timeout(600)
def test_served_mix_with_8192_requests(torch_dev, r2p, synth):
    """Poisson arrivals (200 req/s, 3 s) drawn from the mixed batch's mix plus
    5 % 8192x8192 16-bit radiographs, under an HBM budget: every PNG that
    reaches host memory equals the plain form's PNG of its image."""
    import ctypes as C
    torch = torch_dev
    log("case served mix with 8192^2 requests")
    pool = synth.mixed_batch_device(torch, 32, seed=4, max_side=4096)
    big = synth.radiograph_device(torch, 8192, seed=44)
    tmpl = pool + [big]
    torch.cuda.synchronize()
    opt = r2p.options()
    expect = [encode_plain(torch, r2p, im, t, opt)[0] for im, t in tmpl]
    pins = [t.cpu().pin_memory() for _, t in tmpl]
    rng = np.random.Generator(np.random.PCG64(4))
    lam, secs = 200.0, 3.0
    n = int(lam * secs)
    arr = np.cumsum(rng.exponential(1.0 / lam, n))
    kind = rng.integers(0, len(pool), n).astype(np.int32)
    kind[rng.uniform(size=n) < 0.05] = len(tmpl) - 1
    cfg = r2p.ServerConfig()
    cfg.hbm_budget_bytes = 16 << 30
    cfg.max_batch_images = 64
    cfg.inflight_batches = 3
    cfg.opt = opt
    h = C.c_void_p()
    r2p.check(r2p.lib().r2p_server_create(C.byref(cfg), C.byref(h)))
    try:
        lat, plen, wall, same = r2p.server_loadtest(h, arr, kind, [im for im, _ in tmpl],
                                                    [p.data_ptr() for p in pins], expect)
    finally:
        r2p.lib().r2p_server_destroy(h)
    log(f"served {n} requests ({int((kind == len(tmpl) - 1).sum())} of 8192^2) in {wall:.2f} s, "
        f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms")
    assert np.all(plen == np.array([len(e) for e in expect], np.uint64)[kind])
    bad = np.nonzero(same == 0)[0]
    assert bad.size == 0, f"{bad.size} served PNGs differ from the plain form (first: request {bad[:5]})"
