"""Served form: requests admitted into overlapping device batches produce
the same PNG bytes as each request encoded alone by the plain path, with
fixed- and with dynamic-Huffman blocks."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def plain_png(r2p, img, pixels, opt):
    enc = r2p.Encoder(r2p.filtered_bytes(img), opt)
    out = enc.encode_host(img, pixels)
    enc.close()
    return out


def make_server(r2p, budget=1 << 30, max_batch=8, lanes=3, opt=None):
    cfg = r2p.ServerConfig()
    cfg.hbm_budget_bytes = budget
    cfg.max_batch_images = max_batch
    cfg.inflight_batches = lanes
    cfg.opt = opt or r2p.options()
    h = C.c_void_p()
    r2p.check(r2p.lib().r2p_server_create(C.byref(cfg), C.byref(h)))
    return h, cfg


@pytest.mark.parametrize("huffman", [0, 1])
def test_serve_same_alone_and_in_mix(torch_dev, r2p, synth, huffman):
    log(f"case serve mix huffman={huffman}")
    ims = synth.mixed_batch(20, seed=21, max_side=600, min_side=32)
    h, cfg = make_server(r2p, opt=r2p.options(huffman=huffman))
    L = r2p.lib()
    pins = [np.ascontiguousarray(s.pixels) for s in ims]
    ids = []
    for s, px in zip(ims, pins):
        i = C.c_int64()
        r2p.check(L.r2p_server_submit(h, C.byref(s.img), px.ctypes.data, C.byref(i)))
        ids.append(i.value)
    outs = []
    for i, s in zip(ids, ims):
        cap = r2p.png_bound(s.img)
        buf = np.empty(cap, np.uint8)
        n = C.c_size_t()
        r2p.check(L.r2p_server_wait(h, i, 60000, buf.ctypes.data, cap, C.byref(n)))
        outs.append(buf[: n.value].tobytes())
    st = r2p.ServerStats()
    r2p.check(L.r2p_server_stats_get(h, C.byref(st)))
    assert st.completed == len(ims) and st.batches >= 1
    for s, px, got in zip(ims, pins, outs):
        assert got == plain_png(r2p, s.img, px, cfg.opt)
    # the same request alone through the server
    i = C.c_int64()
    r2p.check(L.r2p_server_submit(h, C.byref(ims[3].img), pins[3].ctypes.data, C.byref(i)))
    cap = r2p.png_bound(ims[3].img)
    buf = np.empty(cap, np.uint8)
    n = C.c_size_t()
    r2p.check(L.r2p_server_wait(h, i.value, 60000, buf.ctypes.data, cap, C.byref(n)))
    assert buf[: n.value].tobytes() == outs[3]
    L.r2p_server_destroy(h)


def test_serve_loadtest_small(torch_dev, r2p, synth):
    log("case serve loadtest")
    ims = synth.mixed_batch(6, seed=22, max_side=512, min_side=64)
    h, cfg = make_server(r2p, budget=256 << 20, max_batch=4, lanes=2)
    L = r2p.lib()
    n = 60
    rng = np.random.default_rng(4)
    arr = np.cumsum(rng.exponential(1.0 / 400.0, n))
    kind = rng.integers(0, len(ims), n).astype(np.int32)
    pins = [np.ascontiguousarray(s.pixels) for s in ims]
    imgs = (r2p.Image * len(ims))(*[s.img for s in ims])
    srcs = (C.c_void_p * len(ims))(*[C.c_void_p(p.ctypes.data) for p in pins])
    lat = np.zeros(n)
    plen = np.zeros(n, np.uint64)
    wall = C.c_double()
    r2p.check(L.r2p_server_loadtest(h, n, arr.ctypes.data_as(C.POINTER(C.c_double)),
                                    kind.ctypes.data_as(C.POINTER(C.c_int)), imgs, srcs,
                                    lat.ctypes.data_as(C.POINTER(C.c_double)),
                                    plen.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(wall), None, None, None))
    assert np.all(lat > 0) and np.all(plen > 0) and wall.value > 0
    alone = [len(plain_png(r2p, s.img, p, cfg.opt)) for s, p in zip(ims, pins)]
    assert all(int(plen[i]) == alone[kind[i]] for i in range(n))
    L.r2p_server_destroy(h)
