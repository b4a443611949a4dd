"""Dynamic-Huffman blocks (r2p_options.huffman = 1): every PNG inflates under
stock zlib to the oracle's filtered bytes, no segment grows past its
fixed-Huffman size, and the batched and served forms reject the option."""
import sys
import zlib

import numpy as np
import pytest

from test_gpu_parity import _check_png, encode_device

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("lz77", [0, 1])
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


def test_dynamic_rejected_outside_plain(torch_dev, r2p):
    log("case dynamic rejected by batch and server")
    with pytest.raises(r2p.R2PError):
        r2p.Batch(4, 1 << 20, 1 << 20, r2p.options(huffman=1))
    b = r2p.Batch(4, 1 << 20, 1 << 20, r2p.options())
    with pytest.raises(r2p.R2PError):
        r2p.check(r2p.lib().r2p_batch_options(b.h, r2p.C.byref(r2p.options(huffman=1))))
    b.close()
