"""Dynamic-Huffman blocks (r2p_options.huffman = 1): every PNG inflates under
stock zlib to the oracle's filtered bytes, no segment grows past its
fixed-Huffman size, and the batched form gives each image the plain path's
bytes."""
import sys
import zlib

import numpy as np
import pytest

from pngutil import check_and_inflate
from test_gpu_batch import plain_png
from test_gpu_parity import _check_png, encode_device

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("lz77", [0, 1, 2])
@pytest.mark.parametrize("kind", ["ct_slice", "radiograph_small", "camera_small"])
def test_dynamic_workloads(torch_dev, r2p, synth, oracle, kind, lz77):
    log(f"case dynamic {kind} lz77={lz77}")
    if kind == "ct_slice":
        s = synth.ct_slice(0)
    elif kind == "radiograph_small":
        s = synth.radiograph(1, size=640)
    else:
        s = synth.camera_rgb(2, 900, 700)
    fixed, _, _, _, _ = encode_device(torch_dev, r2p, s.pixels, s.img, r2p.options(lz77=lz77))
    png, filt, adler, nseg, _ = encode_device(torch_dev, r2p, s.pixels, s.img, r2p.options(lz77=lz77, huffman=1))
    _check_png(oracle, r2p, png, s.pixels, s.img, filt)
    assert adler == zlib.adler32(filt.tobytes())
    assert len(png) < len(fixed), "dynamic blocks should shrink these images"


@pytest.mark.parametrize("seg,slc,w,h,content", [
    (65536, 4096, 2340, 7, "ramp"),         # a 3-byte last slice
    (8192, 4096, 2340, 7, "ramp"),          # three segments, the last a 3-byte slice
    (65536, 16384, 1500, 300, "ramp"),      # segments of four slices, ragged last one
    (1 << 20, 16384, 1500, 300, "ramp"),    # one 1 MiB segment of many slices
    (65536, 16384, 1500, 300, "random"),    # incompressible: the fixed code is kept
    (65536, 16384, 700, 90, "flat"),        # one byte value: a code of two symbols
    (65536, 16384, 1, 1, "flat"),           # a 1x1 image
])
def test_dynamic_geometry_and_content(torch_dev, r2p, oracle, seg, slc, w, h, content):
    log(f"case dynamic seg={seg} slice={slc} {w}x{h} {content}")
    rng = np.random.default_rng(w * 7 + h)
    if content == "ramp":
        px = (np.cumsum(rng.integers(-3, 4, (h, w)), axis=1) % 256).astype(np.uint8)
        px[h // 3: h // 2, :] = 200
    elif content == "random":
        px = rng.integers(0, 256, (h, w), dtype=np.uint8)
    else:
        px = np.full((h, w), 77, np.uint8)
    img = r2p.Image.make(w, h, 1, 8, 0)
    for lz77 in (0, 1):
        base = dict(segment_bytes=seg, slice_bytes=slc, lz77=lz77)
        fixed, _, _, _, _ = encode_device(torch_dev, r2p, px, img, r2p.options(**base))
        png, filt, adler, nseg, _ = encode_device(torch_dev, r2p, px, img, r2p.options(huffman=1, **base))
        _check_png(oracle, r2p, png, px, img, filt)
        assert adler == zlib.adler32(filt.tobytes())
        assert len(png) <= len(fixed)


def _oracle_filtered(oracle, r2p, s):
    p = oracle.DicomParams(*[getattr(s.img.dicom, f) for f, _ in s.img.dicom._fields_])
    flat = s.pixels.reshape(-1)
    if s.img.pix_mode == r2p.PIX_DICOM8:
        ready = oracle.window8(flat, p)
    elif s.img.pix_mode == r2p.PIX_DICOM:
        ready = oracle.dicom_window(flat, p)
    else:
        ready = flat.view(np.uint8)
    rb = r2p.row_bytes(s.img)
    return oracle.filter_rows(ready, s.img.height, rb, rb // s.img.width)


def _first_btype(png):
    """BTYPE of the first DEFLATE block of the PNG's zlib stream."""
    _, z, _ = check_and_inflate(png)
    return (z[2] >> 1) & 3


@pytest.mark.parametrize("graph", [False, True])
def test_dynamic_batch_matches_plain(torch_dev, r2p, synth, oracle, graph):
    """Dynamic blocks in the batched form: each PNG equals the plain path's
    with the image's own options (its slice size included) and inflates to
    the oracle's filtered bytes; a batch switches to fixed blocks and back."""
    torch = torch_dev
    log(f"case dynamic batch graph={graph}")
    ims = synth.mixed_batch(16, seed=13, max_side=900, min_side=16)
    srcs = [torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda() for s in ims]
    fsum = sum(r2p.filtered_bytes(s.img) for s in ims)
    isum = sum(r2p.input_bytes(s.img) for s in ims)
    opt = r2p.options(huffman=1)
    b = r2p.Batch(32, isum, fsum, opt)
    b.set([s.img for s in ims], [t.data_ptr() for t in srcs])
    first = None
    for rep in range(2):
        b.run(graph=graph)
        b.sync()
        got = [b.copy_result(i) for i in range(len(ims))]
        if first is None:
            first = got
            for i, s in enumerate(ims):
                jo = b.job_options(i)
                assert jo.huffman == 1
                assert got[i] == plain_png(r2p, torch, s.img, srcs[i], jo), f"image {i} differs from the plain path"
        else:
            assert got == first
    assert any(_first_btype(png) == 2 for png in first)
    for i in (0, 5, 11):
        s = ims[i]
        _, _, raw = check_and_inflate(first[i])
        assert raw == _oracle_filtered(oracle, r2p, s).tobytes()
    # the same batch with fixed blocks, then dynamic ones again
    b.set_options(r2p.options())
    b.set([s.img for s in ims], [t.data_ptr() for t in srcs])
    b.run(graph=graph)
    b.sync()
    for i in (1, 6):
        assert b.copy_result(i) == plain_png(r2p, torch, ims[i].img, srcs[i], b.job_options(i))
    b.set_options(opt)
    b.set([s.img for s in ims], [t.data_ptr() for t in srcs])
    b.run(graph=graph)
    b.sync()
    assert [b.copy_result(i) for i in range(len(ims))] == first
    b.close()
