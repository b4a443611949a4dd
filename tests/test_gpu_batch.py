"""Batched form: every image's PNG from the batch launch (eager and graph
replays) is byte-identical to the same image encoded alone by the plain
path, and inflates to the oracle's filtered bytes."""
import sys
import zlib

import numpy as np
import pytest

from pngutil import check_and_inflate

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain_png(r2p, torch, img, d_src, opt):
    enc = r2p.Encoder(r2p.filtered_bytes(img), opt)
    cap = r2p.png_bound(img, opt)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
    n, _, _ = enc.sync()
    out = d_png[:n].cpu().numpy().tobytes()
    enc.close()
    return out


def _small_mix(r2p, synth, n, seed):
    ims = synth.mixed_batch(n, seed=seed, max_side=700, min_side=16)
    return ims


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_plain(torch_dev, r2p, synth, oracle, graph):
    torch = torch_dev
    log(f"case batch graph={graph}")
    ims = _small_mix(r2p, synth, 24, seed=5)
    opt = r2p.options()
    srcs = [torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda() for s in ims]
    fsum = sum(r2p.filtered_bytes(s.img) for s in ims)
    isum = sum(r2p.input_bytes(s.img) for s in ims)
    b = r2p.Batch(64, isum, fsum, opt)
    b.set([s.img for s in ims], [t.data_ptr() for t in srcs])
    for rep in range(3):              # replays with eager work and syncs between them
        b.run(graph=graph)
        b.sync()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()
        for i, s in enumerate(ims):
            got = b.copy_result(i)
            if rep == 0:
                exp = plain_png(r2p, torch, s.img, srcs[i], opt)
                assert got == exp, f"image {i} differs from the plain path"
                hdr, z, raw = check_and_inflate(got)
                assert len(raw) == r2p.filtered_bytes(s.img)
            else:
                assert zlib.crc32(got) == zlib.crc32(b.copy_result(i))
    # inflate parity against the oracle for a few images
    for i in (0, 7, 13):
        s = ims[i]
        got = b.copy_result(i)
        hdr, z, raw = check_and_inflate(got)
        p = oracle.DicomParams(*[getattr(s.img.dicom, f) for f, _ in s.img.dicom._fields_])
        flat = s.pixels.reshape(-1)
        if s.img.pix_mode == r2p.PIX_DICOM8:
            ready = oracle.window8(flat, p)
        elif s.img.pix_mode == r2p.PIX_DICOM:
            ready = oracle.dicom_window(flat, p)
        else:
            ready = flat.view(np.uint8)
        rb = r2p.row_bytes(s.img)
        assert raw == oracle.filter_rows(ready, s.img.height, rb, rb // s.img.width).tobytes()
    b.close()


def test_batch_redeclare_and_autotune(torch_dev, r2p, synth):
    torch = torch_dev
    log("case batch redeclare/autotune")
    ims = _small_mix(r2p, synth, 12, seed=9)
    srcs = [torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda() for s in ims]
    fsum = sum(r2p.filtered_bytes(s.img) for s in ims)
    isum = sum(r2p.input_bytes(s.img) for s in ims)
    b = r2p.Batch(32, isum, fsum)
    # declare a subset, run, then the full set: the same graph serves both
    b.set([s.img for s in ims[:5]], [t.data_ptr() for t in srcs[:5]])
    b.run(True); b.sync()
    first = [b.copy_result(i) for i in range(5)]
    b.set([s.img for s in ims], [t.data_ptr() for t in srcs])
    b.run(True); b.sync()
    assert [b.copy_result(i) for i in range(5)] == first
    o, ms = b.autotune()
    assert o.segment_bytes == 65536 and o.wg_threads in (256, 512) and ms > 0
    assert all(c in (0, 4096, 8192, 16384, 32768, 65536) for c in b.class_slices())
    b.run(True); b.sync()
    for i, s in enumerate(ims):
        jo = b.job_options(i)
        assert jo.slice_bytes in (0, 4096, 8192, 16384, 32768, 65536)    # 0: the automatic choice
        assert b.copy_result(i) == plain_png(r2p, torch, s.img, srcs[i], jo)
    b.close()


def test_batch_wide_and_narrow_rows(torch_dev, r2p, oracle):
    """Rows on all sides of the filter's class splits (4096 bytes; 32 KiB,
    above which rows are filtered in tiles), in interleaved order, with ragged
    widths and a 16-bit-stored DICOM image (window table in global memory):
    each PNG equals the plain path's, and the tiled ones inflate to the
    oracle's filtered bytes."""
    torch = torch_dev
    log("case batch wide+narrow+tiled rows")
    rng = np.random.default_rng(11)
    bs16 = r2p.DicomParams.make(16, 16, 15, 1, 1, 1.0, 0.0, 1, 0.0, 3000.0)
    shapes = [(1500, 6, 3, 8, r2p.PIX_READY, None), (300, 9, 1, 8, r2p.PIX_READY, None),
              (17043, 3, 3, 16, r2p.PIX_SWAP16, None), (2100, 5, 1, 16, r2p.PIX_SWAP16, None),
              (1366, 4, 3, 8, r2p.PIX_READY, None), (40000, 2, 1, 8, r2p.PIX_READY, None),
              (1365, 3, 3, 8, r2p.PIX_READY, None), (77, 13, 3, 16, r2p.PIX_SWAP16, None),
              (4096, 3, 1, 8, r2p.PIX_READY, None), (4097, 2, 1, 8, r2p.PIX_READY, None),
              (11000, 4, 3, 8, r2p.PIX_READY, None), (2001, 7, 3, 16, r2p.PIX_SWAP16, None),
              (640, 33, 1, 16, r2p.PIX_DICOM, bs16), (20000, 3, 1, 16, r2p.PIX_DICOM, bs16)]
    imgs, srcs, hosts = [], [], []
    for w, h, ch, bits, mode, dp in shapes:
        n = w * h * ch * (bits // 8)
        px = (np.cumsum(rng.integers(-3, 4, n), dtype=np.int64) % 256).astype(np.uint8)
        px[: n // 4] = rng.integers(0, 256, n // 4, dtype=np.uint8)
        imgs.append(r2p.Image.make(w, h, ch, bits, mode, dp))
        srcs.append(torch.from_numpy(px).cuda())
        hosts.append(px)
    opt = r2p.options()
    fsum = sum(r2p.filtered_bytes(im) for im in imgs)
    isum = sum(r2p.input_bytes(im) for im in imgs)
    b = r2p.Batch(16, isum, fsum, opt)
    b.set(imgs, [t.data_ptr() for t in srcs])
    for graph in (False, True):
        b.run(graph=graph)
        b.sync()
        for i, im in enumerate(imgs):
            got = b.copy_result(i)
            assert got == plain_png(r2p, torch, im, srcs[i], opt), f"image {i} ({im.width}x{im.height}) differs"
    for i, im in enumerate(imgs):
        if r2p.row_bytes(im) <= 32768 and im.pix_mode != r2p.PIX_DICOM:
            continue
        hdr, z, raw = check_and_inflate(b.copy_result(i))
        if im.pix_mode == r2p.PIX_DICOM:
            p = oracle.DicomParams(*[getattr(im.dicom, f) for f, _ in im.dicom._fields_])
            ready = oracle.dicom_window(hosts[i].view("<u2"), p)
        elif im.pix_mode == r2p.PIX_SWAP16:
            ready = hosts[i].view("<u2").byteswap().view(np.uint8)
        else:
            ready = hosts[i]
        rb = r2p.row_bytes(im)
        assert raw == oracle.filter_rows(ready, im.height, rb, rb // im.width).tobytes(), f"image {i} vs oracle"
    b.close()
