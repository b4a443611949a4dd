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
