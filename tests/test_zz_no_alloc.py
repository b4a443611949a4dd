"""The no-allocation plain encode (r2p_encode_device_noalloc): through the
C-ABI its PNGs are byte-identical to r2p_encode_device's, on two calls in a
row (the second with a smaller image, so the shared static scratch holds the
first call's contents), and an image beyond its scratch is refused."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _plain(r2p, torch, img, d_src, opt):
    enc = r2p.Encoder(r2p.filtered_bytes(img), opt)
    cap = r2p.png_bound(img, opt)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
    n, adler, _ = enc.sync()
    out = d_png[:n].cpu().numpy().tobytes()
    enc.close()
    return out, adler


def _noalloc(r2p, torch, img, d_src, opt, d_png, d_res, stream):
    rc = r2p.lib().r2p_encode_device_noalloc(C.byref(img), d_src.data_ptr(), d_png.data_ptr(), d_png.numel(),
                                             d_res.data_ptr(), C.byref(opt) if opt is not None else None,
                                             stream.cuda_stream)
    return rc


def test_noalloc_matches_encoder(torch_dev, r2p, synth):
    torch = torch_dev
    rng = np.random.default_rng(77)
    big = synth.radiograph(1, size=1400)                       # 16-bit float window, 3.9 MB filtered
    small = synth.ct_slice(0, 300)                             # Q15 16 -> 8, 90 KB filtered
    rgb = np.ascontiguousarray((np.cumsum(rng.integers(-2, 3, (257, 333 * 3)), axis=1) % 256).astype(np.uint8))
    cases = [(big.img, big.pixels, None), (small.img, small.pixels, r2p.options(lz77=0)),
             (r2p.Image.make(333, 257, 3, 8, r2p.PIX_READY), rgb, None)]
    stream = torch.cuda.current_stream()
    cap = max(r2p.png_bound(im) for im, _, _ in cases)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    d_res = torch.zeros(2, dtype=torch.int64, device="cuda")
    for im, px, opt in cases:
        log(f"case noalloc {im.width}x{im.height} mode={im.pix_mode}")
        d_src = torch.from_numpy(np.ascontiguousarray(px).view(np.uint8).reshape(-1).copy()).cuda()
        d_png.fill_(0xAB)
        d_res.zero_()
        r2p.check(_noalloc(r2p, torch, im, d_src, opt, d_png, d_res, stream))
        torch.cuda.synchronize()
        n, adler = d_res.cpu().numpy().tolist()
        exp, exp_adler = _plain(r2p, torch, im, d_src, opt or r2p.options())
        assert n == len(exp)
        assert d_png[:n].cpu().numpy().tobytes() == exp, f"{im.width}x{im.height} differs from r2p_encode_device"
        assert (adler & 0xFFFFFFFF) == exp_adler


def test_noalloc_refuses_oversize(torch_dev, r2p):
    torch = torch_dev
    log("case noalloc oversize")
    img = r2p.Image.make(4096, 4096, 1, 8, r2p.PIX_READY)    # 16.8 MB filtered: beyond the scratch
    d_src = torch.zeros(4096 * 4096, dtype=torch.uint8, device="cuda")
    d_png = torch.empty(r2p.png_bound(img), dtype=torch.uint8, device="cuda")
    d_res = torch.zeros(2, dtype=torch.int64, device="cuda")
    rc = _noalloc(r2p, torch, img, d_src, None, d_png, d_res, torch.cuda.current_stream())
    assert rc == r2p.R2P_ERR_SIZE
    assert b"scratch" in r2p.lib().r2p_last_error()
    rc = _noalloc(r2p, torch, img, d_src, r2p.options(huffman=1), d_png, d_res, torch.cuda.current_stream())
    assert rc != 0
