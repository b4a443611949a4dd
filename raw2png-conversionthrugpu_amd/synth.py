"""Seeded synthetic images for the benchmark workloads.

Workloads (numpy PCG64, seed = workload index unless given):
  ct_slice        512x512 uint16, 12 bits stored: background 0, 3-10 overlapping
                  ellipses (centres uniform, semi-axes U[20,200] px, uniform angle,
                  constant level U[800,1600], painter's order), additive N(0,15)
                  rounded, clipped to [0,4095].  DICOM: slope 1, intercept -1024,
                  WC 40, WW 400 -> 8-bit output (Q15 windowing).
  radiograph      3072x3072 uint16, 12 bits stored: 2048 + 1500/8 * sum of 8
                  cosines (|f| <= 0.01 cycles/px, random phase), + 600*y/(H-1),
                  + N(0, sqrt(v)/4), clipped to [0,4095], 10% border = 4095.
                  WC/WW from the 1st/99th percentiles, written as DS strings
                  (two decimals).  16-bit output (float PS3.3 window).
  camera_rgb      6000x4000 RGB8: per-channel linear gradients, 50 flat
                  rectangles (sides U[50,800]), one 1024x1024 area tiled with a
                  random 32x32 texture, N(0,3) rounded, clipped.
  mixed batch     sides floor(exp(U[ln 256, ln 4096])), 50% ct phantom, 30%
                  radiograph background, 20% camera crop.  To keep generation
                  fast, per-image noise is a random crop of one seeded 4096^2
                  noise field per type.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import DicomParams, Image, PIX_DICOM, PIX_DICOM8, PIX_READY


@dataclass
class SynthImage:
    name: str
    pixels: np.ndarray          # uint16 (LE) or uint8, shape (H, W) or (H, W, C)
    img: Image
    meta: dict = field(default_factory=dict)

    @property
    def megapixels(self) -> float:
        return self.img.width * self.img.height / 1e6


def ct_params() -> DicomParams:
    # RescaleSlope "1" / RescaleIntercept "-1024": |intercept| > 1e-9 -> apply_rescale
    # (dicom_loader.cpp:380-388); WC "40" WW "400" -> apply_window (:390-397).
    return DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)


def _ellipses(rng, h, w, out):
    k = int(rng.integers(3, 11))
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    for _ in range(k):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        ry, rx = rng.uniform(20, 200), rng.uniform(20, 200)
        th = rng.uniform(0, math.pi)
        lvl = rng.uniform(800, 1600)
        c, s = math.cos(th), math.sin(th)
        dx, dy = xx - cx, yy - cy
        u = (dx * c + dy * s) / rx
        v = (-dx * s + dy * c) / ry
        out[(u * u + v * v) <= 1.0] = lvl
    return out


def ct_slice(seed: int = 0, size: int = 512, noise: np.ndarray | None = None) -> SynthImage:
    rng = np.random.Generator(np.random.PCG64(seed))
    base = _ellipses(rng, size, size, np.zeros((size, size), np.float32))
    n = noise if noise is not None else rng.standard_normal((size, size), dtype=np.float32)
    v = np.clip(np.rint(base + 15.0 * n), 0, 4095).astype("<u2")
    img = Image.make(size, size, 1, 16, PIX_DICOM8, ct_params())
    return SynthImage("ct_slice", v, img, {"seed": seed})


def _cosine_field(rng, h, w):
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
    acc = np.zeros((h, w), np.float32)
    for _ in range(8):
        f = rng.uniform(0, 0.01)
        ang = rng.uniform(0, 2 * math.pi)
        fx, fy = f * math.cos(ang), f * math.sin(ang)
        ph = rng.uniform(0, 2 * math.pi)
        acc += np.cos(2 * math.pi * (fx * xx + fy * yy) + ph)
    v = 2048.0 + (1500.0 / 8.0) * acc
    v += 600.0 * yy / max(1, h - 1)
    return v


def radiograph(seed: int = 1, size: int = 3072, noise: np.ndarray | None = None) -> SynthImage:
    rng = np.random.Generator(np.random.PCG64(seed))
    v = _cosine_field(rng, size, size)
    n = noise if noise is not None else rng.standard_normal((size, size), dtype=np.float32)
    v = v + np.sqrt(np.maximum(v, 0)) / 4.0 * n
    v = np.clip(np.rint(v), 0, 4095)
    b = int(round(0.10 * size))
    v[:b, :] = 4095; v[-b:, :] = 4095; v[:, :b] = 4095; v[:, -b:] = 4095
    v = v.astype("<u2")
    p1, p99 = np.percentile(v, [1, 99])
    wc = float("%.2f" % ((p1 + p99) / 2.0))
    ww = float("%.2f" % max(p99 - p1, 1.0))
    p = DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 1, wc, ww)
    img = Image.make(size, size, 1, 16, PIX_DICOM, p)
    return SynthImage("radiograph", v, img, {"seed": seed, "wc": wc, "ww": ww})


def camera_rgb(seed: int = 2, width: int = 6000, height: int = 4000,
               noise: np.ndarray | None = None) -> SynthImage:
    rng = np.random.Generator(np.random.PCG64(seed))
    yy, xx = np.mgrid[0:height, 0:width].astype(np.float32)
    img = np.empty((height, width, 3), np.float32)
    for ch in range(3):
        a, bb, c0 = rng.uniform(-120, 120), rng.uniform(-120, 120), rng.uniform(60, 190)
        img[..., ch] = c0 + a * xx / max(1, width - 1) + bb * yy / max(1, height - 1)
    for _ in range(50):
        rh, rw = int(rng.uniform(50, 800)), int(rng.uniform(50, 800))
        y0, x0 = int(rng.integers(0, max(1, height - rh))), int(rng.integers(0, max(1, width - rw)))
        img[y0:y0 + rh, x0:x0 + rw, :] = rng.uniform(0, 255, 3)
    tile = rng.uniform(0, 255, (32, 32, 3)).astype(np.float32)
    th, tw = min(1024, height), min(1024, width)
    ty, tx = int(rng.integers(0, height - th + 1)), int(rng.integers(0, width - tw + 1))
    img[ty:ty + th, tx:tx + tw, :] = np.tile(tile, (th // 32 + 1, tw // 32 + 1, 1))[:th, :tw]
    n = noise if noise is not None else rng.standard_normal((height, width, 3), dtype=np.float32)
    img = np.clip(np.rint(img + 3.0 * n), 0, 255).astype(np.uint8)
    meta = {"seed": seed, "texture": (ty, tx, th, tw)}
    return SynthImage("camera_rgb", img, Image.make(width, height, 3, 8, PIX_READY), meta)


def mixed_batch(n: int = 1024, seed: int = 3, max_side: int = 4096, min_side: int = 256):
    """The mixed-size batch: list of SynthImage (50/30/20 type mix)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    fields = {}

    def field_for(kind, chans):
        if kind not in fields:
            r = np.random.Generator(np.random.PCG64(seed * 1000 + len(fields)))
            shape = (max_side, max_side) if chans == 1 else (max_side, max_side, 3)
            fields[kind] = r.standard_normal(shape, dtype=np.float32)
        return fields[kind]

    out = []
    for i in range(n):
        side = int(math.floor(math.exp(rng.uniform(math.log(min_side), math.log(max_side)))))
        side = max(min_side, min(max_side, side))
        u = rng.uniform()
        oy, ox = (int(rng.integers(0, max_side - side + 1)) for _ in range(2))
        s = int(rng.integers(0, 2**31))
        if u < 0.5:
            f = field_for("ct", 1)[oy:oy + side, ox:ox + side]
            im = ct_slice(s, side, noise=f)
        elif u < 0.8:
            f = field_for("rad", 1)[oy:oy + side, ox:ox + side]
            im = radiograph(s, side, noise=f)
        else:
            f = field_for("rgb", 3)[oy:oy + side, ox:ox + side]
            im = camera_rgb(s, side, side, noise=f)
        im.meta["index"] = i
        out.append(im)
    return out


def workload(name: str) -> SynthImage:
    if name == "ct_slice":
        return ct_slice(0)
    if name == "radiograph":
        return radiograph(1)
    if name == "camera_rgb":
        return camera_rgb(2)
    raise KeyError(name)


# ---------------------------------------------------------------------------
# Device-side generator for the mixed batch (same recipes, drawn with torch's
# CUDA generator so 1024 images of up to 4096^2 appear in seconds).  The
# window of each radiograph uses percentiles of a 1-in-16 pixel subsample.
# ---------------------------------------------------------------------------
def _dev_ct(torch, g, side):
    dev = "cuda"
    yy = torch.arange(side, device=dev, dtype=torch.float32).view(-1, 1)
    xx = torch.arange(side, device=dev, dtype=torch.float32).view(1, -1)
    base = torch.zeros((side, side), device=dev, dtype=torch.float32)
    r = torch.rand(64, generator=g, device=dev).tolist()
    k = 3 + int(r[0] * 8)
    for e in range(k):
        q = r[1 + 6 * e: 7 + 6 * e]
        cy, cx = q[0] * side, q[1] * side
        ry, rx = 20 + 180 * q[2], 20 + 180 * q[3]
        th = q[4] * math.pi
        lvl = 800 + 800 * q[5]
        c, s = math.cos(th), math.sin(th)
        dx, dy = xx - cx, yy - cy
        u = (dx * c + dy * s) / rx
        v = (-dx * s + dy * c) / ry
        base = torch.where(u * u + v * v <= 1.0, torch.full_like(base, lvl), base)
    n = torch.randn((side, side), generator=g, device=dev)
    return torch.clamp(torch.round(base + 15.0 * n), 0, 4095).to(torch.int32)


def _dev_rad(torch, g, side):
    dev = "cuda"
    yy = torch.arange(side, device=dev, dtype=torch.float32).view(-1, 1)
    xx = torch.arange(side, device=dev, dtype=torch.float32).view(1, -1)
    r = torch.rand(24, generator=g, device=dev).tolist()
    acc = torch.zeros((side, side), device=dev, dtype=torch.float32)
    for k in range(8):
        f, ang, ph = 0.01 * r[3 * k], 2 * math.pi * r[3 * k + 1], 2 * math.pi * r[3 * k + 2]
        acc += torch.cos(2 * math.pi * (f * math.cos(ang) * xx + f * math.sin(ang) * yy) + ph)
    v = 2048.0 + (1500.0 / 8.0) * acc + 600.0 * yy / max(1, side - 1)
    v = v + torch.sqrt(torch.clamp(v, min=0)) / 4.0 * torch.randn((side, side), generator=g, device=dev)
    v = torch.clamp(torch.round(v), 0, 4095)
    b = int(round(0.10 * side))
    if b:
        v[:b, :] = 4095; v[-b:, :] = 4095; v[:, :b] = 4095; v[:, -b:] = 4095
    sub = v[::4, ::4].reshape(-1)
    p1, p99 = torch.quantile(sub, torch.tensor([0.01, 0.99], device=dev)).tolist()
    return v.to(torch.int32), float("%.2f" % ((p1 + p99) / 2.0)), float("%.2f" % max(p99 - p1, 1.0))


def _dev_rgb(torch, g, side):
    dev = "cuda"
    yy = torch.arange(side, device=dev, dtype=torch.float32).view(-1, 1, 1)
    xx = torch.arange(side, device=dev, dtype=torch.float32).view(1, -1, 1)
    r = torch.rand(9 + 50 * 7 + 4, generator=g, device=dev).tolist()
    a = torch.tensor([-120 + 240 * r[i] for i in range(3)], device=dev).view(1, 1, 3)
    bb = torch.tensor([-120 + 240 * r[3 + i] for i in range(3)], device=dev).view(1, 1, 3)
    c0 = torch.tensor([60 + 130 * r[6 + i] for i in range(3)], device=dev).view(1, 1, 3)
    img = c0 + a * xx / max(1, side - 1) + bb * yy / max(1, side - 1)
    for e in range(50):
        q = r[9 + 7 * e: 16 + 7 * e]
        rh, rw = int(50 + 750 * q[0]), int(50 + 750 * q[1])
        y0, x0 = int(q[2] * max(1, side - rh)), int(q[3] * max(1, side - rw))
        img[y0:y0 + rh, x0:x0 + rw, :] = torch.tensor([255 * q[4], 255 * q[5], 255 * q[6]], device=dev)
    tile = 255 * torch.rand((32, 32, 3), generator=g, device=dev)
    th = min(1024, side)
    ty = int(r[-2] * (side - th + 1)); tx = int(r[-1] * (side - th + 1))
    img[ty:ty + th, tx:tx + th, :] = tile.repeat(th // 32 + 1, th // 32 + 1, 1)[:th, :th]
    img = img + 3.0 * torch.randn((side, side, 3), generator=g, device=dev)
    return torch.clamp(torch.round(img), 0, 255).to(torch.uint8)


def mixed_batch_device(torch, n: int = 1024, seed: int = 3, max_side: int = 4096, min_side: int = 256):
    """The mixed batch generated in device memory: list of (Image, uint8 tensor)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    g = torch.Generator(device="cuda")
    out = []
    for i in range(n):
        side = int(math.floor(math.exp(rng.uniform(math.log(min_side), math.log(max_side)))))
        side = max(min_side, min(max_side, side))
        u = rng.uniform()
        g.manual_seed(int(rng.integers(0, 2**62)))
        if u < 0.5:
            px = _dev_ct(torch, g, side).to(torch.int16).view(torch.uint8).reshape(-1)
            img = Image.make(side, side, 1, 16, PIX_DICOM8, ct_params())
        elif u < 0.8:
            v, wc, ww = _dev_rad(torch, g, side)
            px = v.to(torch.int16).view(torch.uint8).reshape(-1)
            img = Image.make(side, side, 1, 16, PIX_DICOM, DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 1, wc, ww))
        else:
            px = _dev_rgb(torch, g, side).reshape(-1)
            img = Image.make(side, side, 3, 8, PIX_READY)
        out.append((img, px.contiguous()))
    return out


def radiograph_device(torch, side: int, seed: int):
    """One radiograph-like image (the 3072x3072 workload's generator) of the
    given side, generated in device memory: (Image, uint8 tensor)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    v, wc, ww = _dev_rad(torch, g, side)
    px = v.to(torch.int16).view(torch.uint8).reshape(-1).contiguous()
    img = Image.make(side, side, 1, 16, PIX_DICOM, DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 1, wc, ww))
    return img, px


# ---------------------------------------------------------------------------
# Container writers for the workloads' specified formats: uncompressed
# Explicit VR Little Endian DICOM Part 10 and uncompressed little-endian TIFF.
# ---------------------------------------------------------------------------
def _ds(v) -> bytes:
    s = ("%.6g" % v).encode()
    return s + (b" " if len(s) % 2 else b"")


def write_dicom(path: str, pixels: np.ndarray, p: DicomParams, frames: int = 1,
                slope_str=None, intercept_str=None, wc_str=None, ww_str=None, implicit: bool = False,
                high_bit: bool = True, window=None, rescale=None) -> None:
    """Uncompressed little-endian Part 10 file holding `pixels` (uint16 or
    uint8, (F,)H,W[,C]): Explicit VR, or Implicit VR with `implicit` (the file
    meta group is Explicit VR either way).  `high_bit` False leaves out
    HighBit (0028,0102); `window` / `rescale` write the window or rescale
    attributes (with the *_str values) whatever p's flags say."""
    import struct

    def el_explicit(g, e, vr, val: bytes):
        if len(val) % 2:
            val += b"\0" if vr in (b"UI", b"OB") else b" "
        if vr in (b"OB", b"OW", b"SQ", b"UN", b"UT"):
            return struct.pack("<HH2sHI", g, e, vr, 0, len(val)) + val
        return struct.pack("<HH2sH", g, e, vr, len(val)) + val

    def el(g, e, vr, val: bytes):
        if g == 0x0002 or not implicit:
            return el_explicit(g, e, vr, val)
        if len(val) % 2:
            val += b"\0" if vr in (b"UI", b"OB") else b" "
        return struct.pack("<HHI", g, e, len(val)) + val

    ts = b"1.2.840.10008.1.2" if implicit else b"1.2.840.10008.1.2.1"
    meta_body = el(0x0002, 0x0001, b"OB", b"\0\1") + el(0x0002, 0x0010, b"UI", ts)
    meta = el(0x0002, 0x0000, b"UL", struct.pack("<I", len(meta_body))) + meta_body
    px = np.ascontiguousarray(pixels)
    h, w = px.shape[-2:] if px.ndim == 2 or (px.ndim == 3 and frames > 1) else px.shape[-3:-1]
    spp = 1 if (px.ndim == 2 or (px.ndim == 3 and frames > 1)) else px.shape[-1]
    us = lambda v: struct.pack("<H", v)
    ds = b""
    ds += el(0x0028, 0x0002, b"US", us(spp))
    if frames > 1:
        ds += el(0x0028, 0x0008, b"IS", str(frames).encode())
    ds += el(0x0028, 0x0010, b"US", us(h)) + el(0x0028, 0x0011, b"US", us(w))
    ds += el(0x0028, 0x0100, b"US", us(p.bits_allocated)) + el(0x0028, 0x0101, b"US", us(p.bits_stored))
    if high_bit:
        ds += el(0x0028, 0x0102, b"US", us(p.high_bit))
    ds += el(0x0028, 0x0103, b"US", us(p.pixel_rep))
    if (p.apply_window if window is None else window):
        ds += el(0x0028, 0x1050, b"DS", wc_str or _ds(p.window_center))
        ds += el(0x0028, 0x1051, b"DS", ww_str or _ds(p.window_width))
    if (p.apply_rescale if rescale is None else rescale):
        ds += el(0x0028, 0x1052, b"DS", intercept_str or _ds(p.rescale_intercept))
        ds += el(0x0028, 0x1053, b"DS", slope_str or _ds(p.rescale_slope))
    data = px.astype("<u2" if p.bits_allocated == 16 else np.uint8).tobytes()
    ds += el(0x7FE0, 0x0010, b"OW" if p.bits_allocated == 16 else b"OB", data)
    with open(path, "wb") as f:
        f.write(b"\0" * 128 + b"DICM" + meta + ds)


def write_tiff(path: str, pixels: np.ndarray, rows_per_strip: int = 64, photometric=None) -> None:
    """Uncompressed little-endian TIFF (chunky), uint8 or uint16, gray or RGB;
    PhotometricInterpretation RGB / BlackIsZero unless `photometric` is given
    (0: WhiteIsZero)."""
    import struct
    px = np.ascontiguousarray(pixels)
    h, w = px.shape[:2]
    spp = 1 if px.ndim == 2 else px.shape[2]
    bps = 16 if px.dtype.itemsize == 2 else 8
    data = px.astype("<u2" if bps == 16 else np.uint8).tobytes()
    rb = w * spp * bps // 8
    strips = [(r, min(rows_per_strip, h - r)) for r in range(0, h, rows_per_strip)]
    hdr = 8
    data_off = hdr
    offs = [data_off + r * rb for r, _ in strips]
    cnts = [n * rb for _, n in strips]
    extra_off = data_off + len(data)
    extra = b""

    def arr(vals, typ):
        nonlocal extra
        o = extra_off + len(extra)
        extra += b"".join(struct.pack("<H" if typ == 3 else "<I", v) for v in vals)
        if len(extra) % 2:
            extra += b"\0"
        return o

    entries = []

    def tag(t, typ, vals):
        sz = 2 if typ == 3 else 4
        if len(vals) * sz <= 4:
            raw = b"".join(struct.pack("<H" if typ == 3 else "<I", v) for v in vals).ljust(4, b"\0")
            entries.append(struct.pack("<HHI", t, typ, len(vals)) + raw)
        else:
            entries.append(struct.pack("<HHII", t, typ, len(vals), arr(vals, typ)))

    tag(256, 4, [w]); tag(257, 4, [h]); tag(258, 3, [bps] * spp); tag(259, 3, [1])
    tag(262, 3, [photometric if photometric is not None else (2 if spp >= 3 else 1)])
    tag(273, 4, offs); tag(277, 3, [spp])
    tag(278, 4, [rows_per_strip]); tag(279, 4, cnts); tag(284, 3, [1])
    ifd_off = extra_off + len(extra)
    if ifd_off % 2:
        extra += b"\0"
        ifd_off += 1
    ifd = struct.pack("<H", len(entries)) + b"".join(entries) + struct.pack("<I", 0)
    with open(path, "wb") as f:
        f.write(b"II" + struct.pack("<HI", 42, ifd_off) + data + extra + ifd)
