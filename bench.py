"""Benchmark: encoded MP/s of the MI355X PNG encoder on one GPU.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload radiograph]
                  [--mode plain|batch] [--no-cpu]

A step is one pass of the hot path (window -> filter -> LZ77 + fixed-Huffman
DEFLATE -> checksums -> IDAT assembly) over one batch of synthetic input that
is already resident in HBM; the PNG stays in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import statistics
import sys
import time
import zlib

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(REPO, "raw2png-conversionthrugpu_amd")
METRIC = "encoded MP/s + p50/p99 latency per image, batched vs plain, 1 MI355X, bit-exact"
HBM_PEAK_GBS = 8000.0


def load_pkg():
    name = "raw2png_conversionthrugpu_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def cpu_info():
    """Host cores and CPU model as the OS reports them."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def oracle_encode(oracle, s, level, threads, strip_height=128, want_filtered=False):
    img = s.img
    mode = {0: oracle.MODE_READY, 1: oracle.MODE_SWAP16, 2: oracle.MODE_DICOM, 3: oracle.MODE_DICOM8}[img.pix_mode]
    p = oracle.DicomParams(*[getattr(img.dicom, f) for f, _ in img.dicom._fields_])
    return oracle.encode_png(s.pixels, img.width, img.height, img.channels, img.in_bits, mode, p,
                             strip_height=strip_height, level=level, threads=threads, want_filtered=want_filtered)


def cpu_baseline(oracle, s, level, threads, budget_s=12.0, strip_height=128):
    """Times the oracle's CPU path (window + filter + chunked zlib, the
    reference's deflate_strip scheme) on the same image: 3 warm-ups, then up
    to 15 timed encodes within the budget (bench_deflate's 3 + 15); median."""
    img = s.img
    times, png_len = [], 0
    t_end = time.perf_counter() + budget_s
    it = 0
    while it < 3 or (time.perf_counter() < t_end and it < 18):
        t0 = time.perf_counter()
        e = oracle_encode(oracle, s, level, threads, strip_height)
        dt = time.perf_counter() - t0
        png_len = len(e.png)
        if it >= 3:
            times.append(dt)
        it += 1
    med = statistics.median(times)
    return {"value": round(s.megapixels / med, 2), "unit": "MP/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(times)} timed encodes (after 3 warm-ups) of the same {img.width}x{img.height} "
                      f"image, zlib {oracle.zlib_version()} level {level}, strip height {strip_height}, "
                      f"{threads} threads; median",
            "png_bytes": png_len}


def cpu_leg(args, oracle, synth, r, gpu_pngs):
    """The CPU side: the oracle's filtered bytes of the timed image (the
    bit-exact check of the GPU's PNG), the timed baselines (all usable cores
    and the reference's 6 deflate threads at level 3 on the radiograph; level
    0 on the CT slice, the reference's CPU-runnable case) and zlib's PNG size
    at levels 0, 1 and 3 for the three single-image workloads."""
    out = {"host": cpu_info()}
    # the host's share of cores on a GPU box is 16 whatever nproc says
    threads = min(16, os.cpu_count() or 1)
    out["threads_all"] = threads
    s1 = r["s"]
    e = oracle_encode(oracle, s1, args.cpu_level, threads, want_filtered=True)
    out["bit_exact"] = bool(r["inflated"] == e.filtered.tobytes())
    base = {}
    base["radiograph_level3_all"] = cpu_baseline(oracle, s1, args.cpu_level, threads)
    base["radiograph_level3_6threads"] = cpu_baseline(oracle, s1, args.cpu_level, 6, budget_s=8.0)
    ct = synth.workload("ct_slice")
    base["ct_slice_level0_strip128"] = cpu_baseline(oracle, ct, 0, threads, budget_s=4.0)
    out["baselines"] = {k: {kk: v[kk] for kk in ("value", "cores", "sample")} for k, v in base.items()}
    out["primary"] = base["radiograph_level3_all"]
    ratios = {}
    for name, s in (("ct_slice", ct), ("radiograph", s1), ("camera_rgb", r.get("camera"))):
        if s is None:
            continue
        fb = s.img.height * (s.img.width * s.img.channels * (8 if s.img.pix_mode == 3 else s.img.in_bits) // 8 + 1)
        row = {}
        for lvl in (0, 1, 3):
            row[f"zlib_level{lvl}"] = round(len(oracle_encode(oracle, s, lvl, threads).png) / fb, 4)
        row.update(gpu_pngs.get(name, {}))
        ratios[name] = row
    out["ratios"] = ratios
    return out


def measured_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 byte
    counters (profiles/traffic_*.json, scaled by a copy calibration), when
    they were taken on this workload and configuration; else None."""
    path = os.path.join(REPO, "profiles", "traffic_radiograph3072.json")
    try:
        t = json.load(open(path))
    except (OSError, ValueError):
        return None
    if args.workload != "radiograph" or (t.get("segment_bytes"), t.get("slice_bytes"), t.get("wg_threads"),
                                         t.get("lz77")) != (args.seg, args.slice, args.wg, args.lz77):
        return None
    # counters taken on other kernel sources say nothing about these
    if t.get("kernels_sha16") != kernels_sha16():
        return None
    k = t.get("kernels", {}).get("k_" + kernel)
    return k["bytes"] if k else None


def live_traffic(args):
    """HBM bytes per launch of the plain encode's kernels, measured now: two
    rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE) of a short plain encode
    of this workload, each in a child process that ends with three copies of
    256 MiB; the copies' counters give the bytes per counter unit (FETCH_SIZE
    counts about half the bytes read).  Run before this process touches the
    GPU; None if the profiler is missing or a run fails (bounded by a time
    limit)."""
    import csv
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return None
    child = [sys.executable, "-u", os.path.abspath(__file__), "--mode", "plain", "--no-cpu", "--no-dynamic",
             "--no-ratios", "--steps", "3", "--warmup", "1", "--calib-copy-mib", "256", "--checksum-mib", "0",
             "--traffic", "0", "--workload", args.workload, "--seg", str(args.seg), "--slice", str(args.slice),
             "--wg", str(args.wg), "--lz77", str(args.lz77)]
    env = dict(os.environ, TMPDIR="/tmp")
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="r2p_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "-d", out, "-o", "pmc", "--output-format", "csv", "--"] + child
        log(f"bench: traffic pass {counter}")
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                             start_new_session=True)
        try:
            rc = p.wait(timeout=150)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            log(f"bench: traffic pass {counter} timed out")
            return None
        files = [os.path.join(d, f) for d, _, fs in os.walk(out) for f in fs if f.endswith("counter_collection.csv")]
        if rc != 0 or not files:
            log(f"bench: traffic pass {counter} failed (rc {rc})")
            return None
        vals = {}
        for row in csv.DictReader(open(files[0])):
            if row.get("Counter_Name") == counter:
                vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        shutil.rmtree(out, ignore_errors=True)
        copies = [max(v) for k, v in vals.items() if "copyBuffer" in k]
        if not copies or max(copies) <= 0:
            return None
        unit = float(256 << 20) / (max(copies) * 1024)      # bytes per KiB the counter reports
        for key in ("k_filter", "k_deflate", "k_assemble"):
            v = sorted(x for k, xs in vals.items() if "r2p::" + key in k for x in xs)
            if v:
                got.setdefault(key, {})[counter] = (v[len(v) // 2], unit)
    res = {}
    for key, c in got.items():
        if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
            (f, fu), (w, wu) = c["FETCH_SIZE"], c["WRITE_SIZE"]
            res[key.replace("k_", "")] = {"fetch_kib": round(f), "write_kib": round(w),
                                          "bytes": int(round(fu * 1024 * f + wu * 1024 * w)),
                                          "fetch_factor": round(fu, 4), "write_factor": round(wu, 4)}
    return res or None


def kernels_sha16():
    import hashlib
    with open(os.path.join(PKG_DIR, "csrc", "r2p_kernels.hip"), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def run_plain(args, r2p, synth, torch):
    s = synth.workload(args.workload)
    img = s.img
    opt = r2p.options(lz77=args.lz77, wg_threads=args.wg, segment_bytes=args.seg, slice_bytes=args.slice)
    fb = r2p.filtered_bytes(img)
    enc = r2p.Encoder(fb, opt)
    d_src = torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda()
    cap = r2p.png_bound(img, opt)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    d_filt = torch.empty(fb, dtype=torch.uint8, device="cuda")
    for _ in range(args.warmup):
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
        enc.sync()
    torch.cuda.synchronize()
    lat, ft, dt, st, at = [], [], [], [], []
    # throughput: K encodes issued back to back on the encoder's stream (each
    # a complete per-image encode call; the host does not wait in between)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
    n, adler, nseg = enc.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # latency: K more encodes, each waited for (call to PNG length on the host)
    for _ in range(args.steps):
        ts = time.perf_counter()
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
        n, adler, nseg = enc.sync()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    # per-kernel split (HIP events on the encoder's stream) from the same
    # number of further encodes, outside the timed region: the events cost
    # the stream about 3 us each, so the timed encodes run without them
    enc.set_timing(True)
    for _ in range(args.steps):
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
        enc.sync()
        t = enc.timings()
        ft.append(t.filter_ms); dt.append(t.deflate_ms); st.append(t.scan_ms); at.append(t.assemble_ms)
    # the timed output: its inflated bytes (compared with the oracle's in the
    # CPU leg) and one more encode that also copies the filtered scanlines out
    png = d_png[:n].cpu().numpy().tobytes()
    enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap, d_filt.data_ptr())
    enc.sync()
    assert d_png[:n].cpu().numpy().tobytes() == png, "encode is not deterministic"
    filt = d_filt.cpu().numpy().tobytes()
    z = idat_stream(png)
    inflated = zlib.decompress(z)
    host = run_plain_host(args, r2p, torch, enc, img, s, cap, png)
    mp = s.megapixels
    raw_in = r2p.input_bytes(img)
    kern = {"filter": statistics.mean(ft), "deflate": statistics.mean(dt),
            "scan": statistics.mean(st), "assemble": statistics.mean(at)}
    algo = {"filter": raw_in + fb, "deflate": fb + len(z), "scan": 0, "assemble": 2 * len(z)}
    dom = max(kern, key=kern.get)
    res = {
        "workload": args.workload, "mp": mp, "steps": args.steps, "wall": wall, "lat": lat,
        "png_bytes": n, "filtered_bytes": fb, "raw_in": raw_in, "zlen": len(z),
        "inflated": inflated, "self_consistent": inflated == filt, "host_to_host": host,
        "kern_ms": kern, "algo_bytes": algo, "dominant": dom, "nseg": nseg, "s": s,
    }
    enc.close()
    res["dynamic"] = None if args.no_dynamic else run_plain_dynamic(args, r2p, torch, img, opt, d_src, filt, fb)
    return res


def idat_stream(png: bytes) -> bytes:
    pos, z = 8, []
    while pos < len(png):
        ln = int.from_bytes(png[pos:pos + 4], "big")
        if png[pos + 4:pos + 8] == b"IDAT":
            z.append(png[pos + 8:pos + 8 + ln])
        pos += 12 + ln
    return b"".join(z)


def run_plain_host(args, r2p, torch, enc, img, s, cap, png_dev):
    """The plain form from host memory to host memory (r2p_encode_host: the
    pixels cross PCIe in, the PNG out), pinned buffers on both sides."""
    import ctypes as C
    h_src = torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).pin_memory()
    h_png = torch.empty(cap, dtype=torch.uint8).pin_memory()
    L = r2p.lib()
    n = C.c_size_t()
    for _ in range(3):
        r2p.check(L.r2p_encode_host(enc.h, C.byref(img), h_src.data_ptr(), h_png.data_ptr(), cap, C.byref(n)))
    steps = max(5, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        r2p.check(L.r2p_encode_host(enc.h, C.byref(img), h_src.data_ptr(), h_png.data_ptr(), cap, C.byref(n)))
    dt = (time.perf_counter() - t0) / steps
    same = h_png[:n.value].numpy().tobytes() == png_dev
    return {"mp_per_s": round(s.megapixels / dt, 1), "ms_per_image": round(dt * 1e3, 4),
            "same_png_as_device_form": bool(same),
            "note": "r2p_encode_host: H2D of the pixels, the encode, D2H of the PNG; pinned host buffers"}


def run_ratios(args, r2p, synth, torch):
    """PNG size over filtered bytes of the three single-image workloads with
    fixed-Huffman (the timed form) and dynamic-Huffman blocks, and the time
    per image of each (device-resident input, encode + sync, median of 20
    after 3 warm-ups): dynamic over fixed time."""
    out, samples = {}, {}
    for name in ("ct_slice", "radiograph", "camera_rgb"):
        s = synth.workload(name)
        samples[name] = s
        d_src = torch.from_numpy(np.ascontiguousarray(s.pixels).view(np.uint8).reshape(-1).copy()).cuda()
        fb = r2p.filtered_bytes(s.img)
        row, ms = {}, {}
        for label, huff in (("gpu_fixed", 0), ("gpu_dynamic", 1)):
            o = r2p.options(huffman=huff)
            enc = r2p.Encoder(fb, o)
            cap = r2p.png_bound(s.img, o)
            d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
            ts = []
            for k in range(23):
                t0 = time.perf_counter()
                enc.encode_device(s.img, d_src.data_ptr(), d_png.data_ptr(), cap)
                n, _, _ = enc.sync()
                if k >= 3:
                    ts.append(time.perf_counter() - t0)
            enc.close()
            row[label] = round(n / fb, 4)
            ms[label] = float(np.median(ts)) * 1e3
        row["gpu_ms_fixed"] = round(ms["gpu_fixed"], 4)
        row["gpu_ms_dynamic"] = round(ms["gpu_dynamic"], 4)
        row["dynamic_over_fixed_time"] = round(ms["gpu_dynamic"] / ms["gpu_fixed"], 3)
        out[name] = row
        del d_src
    return out, samples


def run_plain_dynamic(args, r2p, torch, img, opt, d_src, filt, fb):
    """The same image with a dynamic-Huffman block per segment (huffman=1):
    PNG size and time per image, outside the timed region of the headline."""
    o = r2p.options(lz77=opt.lz77, wg_threads=opt.wg_threads, segment_bytes=opt.segment_bytes,
                    slice_bytes=opt.slice_bytes, huffman=1)
    enc = r2p.Encoder(fb, o)
    cap = r2p.png_bound(img, o)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
        enc.sync()
    steps = max(5, min(args.steps, 50))
    t0 = time.perf_counter()
    for _ in range(steps):
        enc.encode_device(img, d_src.data_ptr(), d_png.data_ptr(), cap)
        n, _, _ = enc.sync()
    ms = (time.perf_counter() - t0) / steps * 1e3
    png = d_png[:n].cpu().numpy().tobytes()
    enc.close()
    z = idat_stream(png)
    return {"png_over_filtered": round(n / fb, 4), "ms_per_image": round(ms, 4),
            "mp_per_s": round(img.width * img.height / 1e6 / (ms * 1e-3), 1),
            "inflates_to_filtered": zlib.decompress(z) == filt}


def run_checksums(args, r2p, torch):
    """The standalone Adler-32 / CRC-32 kernels (per-block partials, then one
    combine workgroup) over a large device buffer of random bytes; the HBM
    rate is the bytes read over the time of both launches (HIP events on
    the stream they run on)."""
    n = args.checksum_mib << 20
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    d = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda", generator=g)
    work = torch.empty(r2p.checksum_work_bytes(n), dtype=torch.uint8, device="cuda")
    res = torch.zeros(2, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    out = {"bytes": n, "data": "random bytes, device-resident"}
    for name, fn, off in (("adler32", r2p.adler32_device, 0), ("crc32", r2p.crc32_device, 8)):
        for _ in range(3):
            fn(d.data_ptr(), n, work.data_ptr(), res.data_ptr() + off, sp)
        ts = []
        for _ in range(10):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn(d.data_ptr(), n, work.data_ptr(), res.data_ptr() + off, sp)
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        ms = statistics.median(ts)
        gbs = n / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "gb_per_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    torch.cuda.synchronize()
    v = res.cpu().numpy().view(np.uint32)
    host = d.cpu().numpy().tobytes()
    out["zlib_equal"] = bool(v[0] == zlib.adler32(host) and v[2] == zlib.crc32(host))
    log("bench: checksums " + json.dumps(out))
    return out


def run_batch(args, r2p, synth, torch):
    """Mixed batch: the batched form (graph replays) against the plain form on the same images."""
    t0 = time.perf_counter()
    items = synth.mixed_batch_device(torch, args.batch_n, seed=3, max_side=args.batch_max_side)
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    imgs = [im for im, _ in items]
    srcs = [t.data_ptr() for _, t in items]
    mp = sum(im.width * im.height for im in imgs) / 1e6
    fsum = sum(r2p.filtered_bytes(im) for im in imgs)
    isum = sum(r2p.input_bytes(im) for im in imgs)
    log(f"bench: batch of {len(imgs)} images, {mp:.1f} MP, {fsum / 1e9:.2f} GB filtered (gen {gen_s:.1f}s)")
    opt = r2p.options()
    b = r2p.Batch(len(imgs), isum, fsum, opt)
    b.set(imgs, srcs)
    tuned, tuned_ms = b.autotune() if args.autotune else (b.opt, None)
    for _ in range(2):
        b.run(True); b.sync()
    steps = max(3, args.batch_steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gtimes = []
    for _ in range(steps):
        b.run(True); b.sync()
        gtimes.append(b.timings().total_ms)
    torch.cuda.synchronize()
    bwall = (time.perf_counter() - t1) / steps
    # per-kernel split from eager (non-graph) runs of the same batch
    ktimes = []
    for _ in range(2):
        b.run(False); b.sync()
        t = b.timings()
        ktimes.append((t.filter_ms, t.deflate_ms, t.scan_ms, t.assemble_ms, t.total_ms))
    # plain form, one image at a time, with the options each image got in the
    # batch (the tuned slice size of its size class): one encoder per slice size
    maxf = max(r2p.filtered_bytes(im) for im in imgs)
    job_opt = [b.job_options(i) for i in range(len(imgs))]
    encs = {}
    for jo in job_opt:
        if jo.slice_bytes not in encs:
            encs[jo.slice_bytes] = r2p.Encoder(maxf, jo)
    cap = max(r2p.png_bound(im, tuned) for im in imgs)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    for i, (im, sp) in list(enumerate(zip(imgs, srcs)))[:8]:
        enc = encs[job_opt[i].slice_bytes]
        enc.encode_device(im, sp, d_png.data_ptr(), cap); enc.sync()
    lat = []
    differ = 0
    torch.cuda.synchronize()
    for i, (im, sp) in enumerate(zip(imgs, srcs)):
        enc = encs[job_opt[i].slice_bytes]
        ts = time.perf_counter()
        enc.encode_device(im, sp, d_png.data_ptr(), cap)
        n, _, _ = enc.sync()
        lat.append(time.perf_counter() - ts)
        # every image's PNG against the batch's (the copies are not timed)
        differ += d_png[:n].cpu().numpy().tobytes() != b.copy_result(i)
    pwall = sum(lat)
    # the plain form's calls issued back to back (no wait per image; where the
    # next image uses another slice size, and so another encoder and stream,
    # the previous encoder is waited for, so that the encodes stay one at a time)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    prev = None
    for i, (im, sp) in enumerate(zip(imgs, srcs)):
        enc = encs[job_opt[i].slice_bytes]
        if prev is not None and prev is not enc:
            prev.sync()
        enc.encode_device(im, sp, d_png.data_ptr(), cap)
        prev = enc
    prev.sync()
    torch.cuda.synchronize()
    pwall_b2b = time.perf_counter() - t2
    class_slices = b.class_slices()
    for enc in encs.values():
        enc.close()
    png_total = sum(b.result(i)[1] for i in range(len(imgs)))
    b.close()
    dyn = run_batch_dynamic(args, r2p, torch, imgs, srcs, isum, fsum, mp, tuned)
    kt = np.array(ktimes).mean(axis=0)
    lat_ms = [x * 1e3 for x in lat]
    return {
        "images": len(imgs), "megapixels": round(mp, 1), "filtered_gb": round(fsum / 1e9, 3),
        "batched_mp_per_s": round(mp / bwall, 1), "batched_ms": round(bwall * 1e3, 3),
        "graph_device_ms": round(float(np.median(gtimes)), 3),
        "plain_mp_per_s": round(mp / pwall, 1), "plain_ms": round(pwall * 1e3, 3),
        "speedup_batched_over_plain": round(pwall / bwall, 3),
        "plain_back_to_back_mp_per_s": round(mp / pwall_b2b, 1),
        "speedup_batched_over_plain_back_to_back": round(pwall_b2b / bwall, 3),
        "plain_latency_ms": {"p50": round(pct(lat_ms, 50), 4), "p99": round(pct(lat_ms, 99), 4)},
        "batched_latency_ms": {"p50": round(bwall * 1e3, 3), "p99": round(bwall * 1e3, 3)},
        "kernels_ms": {"filter": round(kt[0], 3), "deflate": round(kt[1], 3), "scan": round(kt[2], 3),
                       "assemble": round(kt[3], 3)},
        "filter_gbs": round((isum + fsum) / (kt[0] * 1e-3) / 1e9, 1),
        "deflate_gbs": round((fsum + png_total) / (kt[1] * 1e-3) / 1e9, 1),
        "autotuned": {"segment_bytes": tuned.segment_bytes,
                      "slice_bytes_by_class": dict(zip(("under_2MiB", "under_8MiB", "larger"), class_slices)),
                      "wg_threads": tuned.wg_threads,
                      "band_rows": tuned.band_rows, "best_ms": tuned_ms},
        "png_over_filtered": round(png_total / fsum, 4),
        "byte_identical_to_plain": differ == 0, "images_compared": len(imgs), "steps": steps,
        "dynamic_huffman": dyn,
    }


def run_batch_dynamic(args, r2p, torch, imgs, srcs, isum, fsum, mp, tuned):
    """The same batch with dynamic-Huffman blocks (graph replays, slices by
    image size), and every image's PNG compared with the plain form's."""
    o = r2p.options(huffman=1, band_rows=tuned.band_rows)
    b = r2p.Batch(len(imgs), isum, fsum, o)
    b.set(imgs, srcs)
    for _ in range(2):
        b.run(True); b.sync()
    steps = max(3, args.batch_steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        b.run(True); b.sync()
    torch.cuda.synchronize()
    bwall = (time.perf_counter() - t1) / steps
    job_opt = [b.job_options(i) for i in range(len(imgs))]
    maxf = max(r2p.filtered_bytes(im) for im in imgs)
    encs = {}
    for jo in job_opt:
        if jo.slice_bytes not in encs:
            encs[jo.slice_bytes] = r2p.Encoder(maxf, jo)
    cap = max(r2p.png_bound(im, o) for im in imgs)
    d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
    differ = 0
    for i, (im, sp) in enumerate(zip(imgs, srcs)):
        enc = encs[job_opt[i].slice_bytes]
        enc.encode_device(im, sp, d_png.data_ptr(), cap)
        n, _, _ = enc.sync()
        differ += d_png[:n].cpu().numpy().tobytes() != b.copy_result(i)
    for enc in encs.values():
        enc.close()
    png_total = sum(b.result(i)[1] for i in range(len(imgs)))
    b.close()
    return {"batched_mp_per_s": round(mp / bwall, 1), "batched_ms": round(bwall * 1e3, 3),
            "png_over_filtered": round(png_total / fsum, 4), "byte_identical_to_plain": differ == 0,
            "images_compared": len(imgs), "steps": steps}


def run_serve(args, r2p, synth, torch):
    """Served form: open-loop Poisson arrivals drawn from the mixed batch's
    size/type mix plus 5% 8192x8192 uint16 radiographs, admitted against a
    64 GB HBM budget; latency is arrival -> PNG bytes in host memory.  Every
    request's PNG is compared byte for byte with the plain form's PNG of its
    image (native verifier threads, after the latency is taken)."""
    import ctypes as C
    L = r2p.lib()
    rng = np.random.Generator(np.random.PCG64(4))
    pool = synth.mixed_batch_device(torch, args.serve_pool, seed=4, max_side=args.batch_max_side)
    big = synth.radiograph_device(torch, 8192, seed=44)
    tmpl = pool + [big]
    torch.cuda.synchronize()
    pins = [t.cpu().pin_memory() for _, t in tmpl]
    imgs = [im for im, _ in tmpl]
    mp_of = np.array([im.width * im.height / 1e6 for im in imgs])
    in_of = np.array([r2p.input_bytes(im) for im in imgs], dtype=np.float64)

    def expected(opt):
        """The plain form's PNG of every template."""
        exp = []
        for im, t in tmpl:
            enc = r2p.Encoder(r2p.filtered_bytes(im), opt)
            cap = r2p.png_bound(im, opt)
            d_png = torch.empty(cap, dtype=torch.uint8, device="cuda")
            enc.encode_device(im, t.data_ptr(), d_png.data_ptr(), cap)
            n, _, _ = enc.sync()
            enc.close()
            exp.append(d_png[:n].cpu().numpy().tobytes())
        return exp

    def one_rate(lam, seconds, opt, expect):
        n = int(lam * seconds)
        arr = np.cumsum(rng.exponential(1.0 / lam, n))
        kind = rng.integers(0, len(pool), n).astype(np.int32)
        kind[rng.uniform(size=n) < 0.05] = len(tmpl) - 1
        cfg = r2p.ServerConfig()
        cfg.hbm_budget_bytes = 64 << 30
        cfg.max_batch_images = args.serve_batch
        cfg.inflight_batches = args.serve_lanes
        cfg.opt = opt
        h = C.c_void_p()
        r2p.check(L.r2p_server_create(C.byref(cfg), C.byref(h)))
        try:
            lat, plen, wall, same = r2p.server_loadtest(h, arr, kind, imgs, [p.data_ptr() for p in pins], expect)
            st = r2p.ServerStats()
            r2p.check(L.r2p_server_stats_get(h, C.byref(st)))
        finally:
            L.r2p_server_destroy(h)
        mp = float(mp_of[kind].sum())
        return {
            "requests": n, "offered_mp_per_s": round(mp / float(arr[-1]), 1),
            "mp_per_s": round(mp / wall, 1), "wall_s": round(wall, 3),
            "h2d_gb_per_s": round(float(in_of[kind].sum()) / wall / 1e9, 2),
            "d2h_gb_per_s": round(float(plen.astype(np.float64).sum()) / wall / 1e9, 2),
            "latency_ms": {q: round(float(np.percentile(lat, int(q[1:]))) * 1e3, 3) for q in ("p50", "p95", "p99")},
            "batches": int(st.batches), "mean_batch_images": round(st.mean_batch_images, 2),
            "peak_admitted_gb": round(st.peak_admitted_bytes / 1e9, 3),
            "png_over_input": round(float(plen.astype(np.float64).sum()) / float(in_of[kind].sum()), 4),
            "png_bytes_equal_plain": int(same.sum()) == n, "requests_compared": n,
        }

    out = {"seconds_per_rate": args.serve_seconds, "hbm_budget_gb": 64, "lanes": args.serve_lanes,
           "max_batch_images": args.serve_batch, "pool": len(pool), "big_requests": "5% 8192x8192 16-bit"}
    opt = r2p.options()
    expect = expected(opt)
    for lam in args.serve_rates:
        out[str(lam)] = one_rate(lam, args.serve_seconds, opt, expect)
        log(f"bench: serve lambda={lam}: {json.dumps(out[str(lam)])}")
    if args.serve_dynamic_seconds > 0:
        # the same mix with dynamic-Huffman blocks, at the middle rate
        dopt = r2p.options(huffman=1)
        lam = args.serve_rates[len(args.serve_rates) // 2]
        out["dynamic_huffman"] = {str(lam): one_rate(lam, args.serve_dynamic_seconds, dopt, expected(dopt)),
                                  "seconds": args.serve_dynamic_seconds}
        log(f"bench: serve dynamic lambda={lam}: {json.dumps(out['dynamic_huffman'])}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="radiograph", choices=["ct_slice", "radiograph", "camera_rgb"])
    ap.add_argument("--mode", default="all", choices=["plain", "batch", "serve", "all"])
    ap.add_argument("--batch-n", type=int, default=1024)
    ap.add_argument("--batch-max-side", type=int, default=4096)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--lz77", type=int, default=1)
    ap.add_argument("--wg", type=int, default=256, choices=[256, 512])
    ap.add_argument("--seg", type=int, default=65536, choices=[16384, 32768, 65536])
    ap.add_argument("--slice", type=int, default=0, choices=[0, 4096, 8192, 16384, 32768, 65536],
                    help="DEFLATE slice bytes (0: by image size; 16 KiB for the 3072x3072 radiograph)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic-Huffman encodes of the plain workload")
    ap.add_argument("--no-ratios", action="store_true",
                    help="skip the fixed/dynamic size and time table of the three single-image workloads")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[200.0, 800.0, 2000.0])
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window per rate")
    ap.add_argument("--serve-dynamic-seconds", type=float, default=20.0,
                    help="arrival window of the dynamic-Huffman served run (0: skip)")
    ap.add_argument("--serve-pool", type=int, default=64, help="distinct request images of the mix")
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--traffic", type=int, default=1,
                    help="measure the plain kernels' HBM bytes with two rocprofv3 --pmc child runs first (0: no)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, copy this many MiB device-to-device three times (profiler calibration)")
    ap.add_argument("--cpu-level", type=int, default=3)
    ap.add_argument("--checksum-mib", type=int, default=1024, help="buffer for the standalone checksum kernels (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this encoder is measured on one GPU only; --gpus must be 1 "
              "(the path shards by image with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    # the byte counters first, while this process has not touched the GPU
    traffic = live_traffic(args) if args.traffic else None
    import torch
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    r2p = load_pkg()
    import importlib
    synth = importlib.import_module("raw2png_conversionthrugpu_amd.synth")
    torch.cuda.set_device(0)
    log(f"bench: workload={args.workload} steps={args.steps} warmup={args.warmup}")
    r = run_plain(args, r2p, synth, torch)
    kern = r["kern_ms"]
    dom = r["dominant"]
    achieved = r["algo_bytes"][dom] / (kern[dom] * 1e-3) / 1e9
    value = r["mp"] * r["steps"] / r["wall"]
    lat_ms = [x * 1e3 for x in r["lat"]]
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "MP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r["wall"] / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{args.workload} {r['s'].img.width}x{r['s'].img.height}",
                   "form": "plain per-image", "segment_bytes": args.seg, "slice_bytes": args.slice or "auto (16384 here)",
                   "wg_threads": args.wg, "lz77": args.lz77,
                   "bit_exact": None},
        "timing": "value: K per-image encode calls issued back to back, synchronized before and after; "
                  "latency_ms: K more calls, each waited for",
        "latency_ms": {"p50": round(pct(lat_ms, 50), 4), "p99": round(pct(lat_ms, 99), 4)},
        "mp_per_s_each_waited_for": round(r["mp"] / (sum(lat_ms) / len(lat_ms) * 1e-3), 2),
        "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
        "ratio": {"gpu_png_over_filtered": round(r["png_bytes"] / r["filtered_bytes"], 4)},
        "host_to_host": r["host_to_host"],
        "dynamic_huffman": r["dynamic"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None},
    }
    if traffic and dom in traffic:
        out["roofline"]["traffic"] = traffic[dom]["bytes"]
        out["roofline"]["traffic_source"] = "measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, copy-calibrated"
        out["traffic_by_kernel"] = {k: dict(v, algorithmic=int(r["algo_bytes"][k])) for k, v in traffic.items()
                                    if k in r["algo_bytes"]}
    else:
        out["roofline"]["traffic"] = measured_traffic(args, dom)
        if out["roofline"]["traffic"] is not None:
            out["roofline"]["traffic_source"] = "profiles/traffic_radiograph3072.json (counters of these kernel sources)"
    if r["dynamic"]:
        out["ratio"]["gpu_dynamic_huffman_png_over_filtered"] = r["dynamic"]["png_over_filtered"]
    log("bench: plain " + json.dumps({k: out[k] for k in ("value", "ms_per_step", "latency_ms", "kernels_ms", "roofline")}))
    fk = "filter"
    fach = r["algo_bytes"][fk] / (kern[fk] * 1e-3) / 1e9
    out["filter_kernel"] = {"achieved": round(fach, 1), "frac": round(fach / HBM_PEAK_GBS, 4)}
    if args.no_ratios:
        gpu_ratios, r["camera"] = {}, None
    else:
        gpu_ratios, samples = run_ratios(args, r2p, synth, torch)
        r["camera"] = samples["camera_rgb"]
    if not args.no_cpu:
        # the oracle is the checker of the timed PNG and the timed CPU baseline
        import oracle as O
        c = cpu_leg(args, O, synth, r, gpu_ratios)
        out["config"]["bit_exact"] = c["bit_exact"]
        cb = c["primary"]
        out["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        out["cpu_baseline"]["host"] = c["host"]
        out["cpu_baselines"] = c["baselines"]
        out["ratio"]["by_workload"] = c["ratios"]
        log("bench: cpu " + json.dumps({k: c[k] for k in ("bit_exact", "host", "baselines")}))
    else:
        out["ratio"]["by_workload"] = gpu_ratios
    if args.checksum_mib > 0:
        out["checksum_kernels"] = run_checksums(args, r2p, torch)
    if args.mode in ("batch", "all"):
        out["batched"] = run_batch(args, r2p, synth, torch)
    if args.mode in ("serve", "all"):
        out["served"] = run_serve(args, r2p, synth, torch)
    if args.calib_copy_mib:
        # a plain device copy of known size, for scaling the HBM byte counters
        # of a profiler run (FETCH_SIZE / WRITE_SIZE) to bytes
        n = args.calib_copy_mib << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"bench: calibration copy of {n} bytes x3 done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
