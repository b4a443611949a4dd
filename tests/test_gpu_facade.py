"""The drop-in C++ API (reference header signatures) on the GPU: runs the
tests/cpp/facade_check driver built by build()."""
import os
import subprocess
import sys

import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu


def test_facade_check(torch_dev, tmp_path):
    exe = os.path.join(REPO, "tests", "cpp", "facade_check")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/facade_check was not built (run __graft_entry__.build())")
    print("case facade_check", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=110)
    sys.stdout.write(r.stdout)
    sys.stderr.write(r.stderr)
    assert r.returncode == 0 and "ALL OK" in r.stdout


@pytest.mark.parametrize("gpu_deflate", [0, 1])
def test_file_readers_through_pipeline(torch_dev, r2p, synth, oracle, tmp_path, gpu_deflate):
    """TIFF and DICOM Part 10 files -> encode_*_to_png -> PIL / oracle."""
    import io
    import numpy as np
    from PIL import Image as PILImage
    exe = os.path.join(REPO, "tests", "cpp", "facade_check")
    print(f"case readers gpu_deflate={gpu_deflate}", flush=True)
    # RGB8 TIFF
    cam = synth.camera_rgb(2, 300, 200)
    tif = str(tmp_path / "cam.tif")
    synth.write_tiff(tif, cam.pixels, rows_per_strip=37)
    out = str(tmp_path / "cam.png")
    r = subprocess.run([exe, "encode", "tiff", tif, out, str(gpu_deflate)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(np.array(PILImage.open(out)), cam.pixels)
    # 16-bit grayscale TIFF (loader byte-swaps to big-endian)
    g16 = np.random.default_rng(1).integers(0, 65536, (50, 61)).astype(np.uint16)
    synth.write_tiff(str(tmp_path / "g.tif"), g16, rows_per_strip=7)
    r = subprocess.run([exe, "encode", "tiff", str(tmp_path / "g.tif"), str(tmp_path / "g.png"), str(gpu_deflate)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(np.array(PILImage.open(str(tmp_path / "g.png"))), g16)
    # CT slice DICOM: the single-file path keeps 16 bits (float window, gpu_filter.cu:71-162)
    ct = synth.ct_slice(0, 128)
    dcm = str(tmp_path / "ct.dcm")
    synth.write_dicom(dcm, ct.pixels, ct.img.dicom)
    r = subprocess.run([exe, "encode", "dicom", dcm, str(tmp_path / "ct.png"), str(gpu_deflate)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    p = oracle.DicomParams(*[getattr(ct.img.dicom, f) for f, _ in ct.img.dicom._fields_])
    exp = oracle.dicom_window(ct.pixels.reshape(-1), p).view(">u2").reshape(128, 128)
    assert np.array_equal(np.array(PILImage.open(str(tmp_path / "ct.png"))).astype(np.int64), exp.astype(np.int64))
    # multi-frame: more frames than the output stage keeps in flight
    nfr = 7
    frames = np.stack([synth.ct_slice(k, 64).pixels for k in range(nfr)])
    synth.write_dicom(str(tmp_path / "mf.dcm"), frames, ct.img.dicom, frames=nfr)
    os.makedirs(str(tmp_path / "mf"), exist_ok=True)
    r = subprocess.run([exe, "encode", "dicom-all", str(tmp_path / "mf.dcm"), str(tmp_path / "mf"), str(gpu_deflate)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    for k in range(nfr):
        got = np.array(PILImage.open(str(tmp_path / "mf" / f"slice_{k:04d}.png"))).astype(np.int64)
        exp = oracle.dicom_window(frames[k].reshape(-1), p).view(">u2").reshape(64, 64)
        assert np.array_equal(got, exp.astype(np.int64))
    # the output folder is created, parents included (pipeline.cpp:1768-1774) ...
    r = subprocess.run([exe, "encode", "dicom-all", str(tmp_path / "mf.dcm"), str(tmp_path / "new" / "sub"),
                        str(gpu_deflate)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(os.listdir(str(tmp_path / "new" / "sub"))) == [f"slice_{k:04d}.png" for k in range(nfr)]
    # ... and a folder that cannot be created fails the call
    open(str(tmp_path / "plainfile"), "w").write("x")
    r = subprocess.run([exe, "encode", "dicom-all", str(tmp_path / "mf.dcm"), str(tmp_path / "plainfile" / "sub"),
                        str(gpu_deflate)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    # PipelineConfig::frame selects one frame of the multi-frame file
    for k in (0, 4):
        out = str(tmp_path / f"frame{k}.png")
        r = subprocess.run([exe, "encode", "dicom-frame", str(tmp_path / "mf.dcm"), out, str(gpu_deflate), str(k)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        exp = oracle.dicom_window(frames[k].reshape(-1), p).view(">u2").reshape(64, 64)
        assert np.array_equal(np.array(PILImage.open(out)).astype(np.int64), exp.astype(np.int64))
    r = subprocess.run([exe, "encode", "dicom-frame", str(tmp_path / "mf.dcm"), str(tmp_path / "bad.png"),
                        str(gpu_deflate), str(nfr)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0                       # frame out of range
    # Implicit VR little endian, same pixels and attributes
    synth.write_dicom(str(tmp_path / "ct_ivr.dcm"), ct.pixels, ct.img.dicom, implicit=True)
    r = subprocess.run([exe, "encode", "dicom", str(tmp_path / "ct_ivr.dcm"), str(tmp_path / "ct_ivr.png"),
                        str(gpu_deflate)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert open(str(tmp_path / "ct_ivr.png"), "rb").read() == open(str(tmp_path / "ct.png"), "rb").read()
    # RGB DICOM (8-bit, 3 samples per pixel; the 8-bit branch of the window,
    # gpu_filter.cu:122-130), with a rescale so the window is not the identity
    rgb = synth.camera_rgb(5, 97, 61).pixels
    prgb = r2p.DicomParams.make(8, 8, 7, 0, 1, 0.5, 10.0, 0)
    synth.write_dicom(str(tmp_path / "rgb.dcm"), rgb, prgb)
    r = subprocess.run([exe, "encode", "dicom", str(tmp_path / "rgb.dcm"), str(tmp_path / "rgb.png"), str(gpu_deflate)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    op = oracle.DicomParams(*[getattr(prgb, f) for f, _ in prgb._fields_])
    exp = oracle.dicom_window(rgb.reshape(-1), op).reshape(61, 97, 3)
    got = np.array(PILImage.open(str(tmp_path / "rgb.png")))
    assert got.shape == (61, 97, 3) and np.array_equal(got, exp)


def test_reader_attribute_edges(torch_dev, r2p, synth, oracle, tmp_path):
    """Reader cases the loaders decide, through encode_*_to_png: a
    WhiteIsZero TIFF is inverted (image_loader.cpp:204-210); DICOM without
    HighBit takes BitsStored - 1 (dicom_loader.cpp:329-330); WindowWidth 0
    turns the window off and RescaleSlope "1" / RescaleIntercept "0" turn the
    rescale off (dicom_loader.cpp:380-397)."""
    import numpy as np
    from PIL import Image as PILImage
    exe = os.path.join(REPO, "tests", "cpp", "facade_check")
    print("case reader attribute edges", flush=True)

    def encode(kind, src, name):
        out = str(tmp_path / name)
        r = subprocess.run([exe, "encode", kind, src, out, "1"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        return np.array(PILImage.open(out)).astype(np.int64)

    rng = np.random.default_rng(5)
    g8 = rng.integers(0, 256, (30, 41)).astype(np.uint8)
    synth.write_tiff(str(tmp_path / "w8.tif"), g8, rows_per_strip=7, photometric=0)
    assert np.array_equal(encode("tiff", str(tmp_path / "w8.tif"), "w8.png"), 255 - g8.astype(np.int64))
    g16 = rng.integers(0, 65536, (20, 33)).astype(np.uint16)
    synth.write_tiff(str(tmp_path / "w16.tif"), g16, rows_per_strip=6, photometric=0)
    assert np.array_equal(encode("tiff", str(tmp_path / "w16.tif"), "w16.png"), 65535 - g16.astype(np.int64))

    ct = synth.ct_slice(3, 96)
    px = ct.pixels

    def expect(p):
        op = oracle.DicomParams(*[getattr(p, f) for f, _ in p._fields_])
        return oracle.dicom_window(px.reshape(-1), op).view(">u2").reshape(px.shape).astype(np.int64)

    # HighBit absent: BitsStored - 1 (here 11, where a default of 15 would shift the codes)
    p = r2p.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)
    synth.write_dicom(str(tmp_path / "nohb.dcm"), px, p, high_bit=False)
    assert np.array_equal(encode("dicom", str(tmp_path / "nohb.dcm"), "nohb.png"), expect(p))
    # WindowWidth 0: no window (the rescale still applies)
    synth.write_dicom(str(tmp_path / "ww0.dcm"), px, p, ww_str=b"0")
    p_nowin = r2p.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 0, 0.0, 0.0)
    got = encode("dicom", str(tmp_path / "ww0.dcm"), "ww0.png")
    assert np.array_equal(got, expect(p_nowin)) and not np.array_equal(got, expect(p))
    # RescaleSlope "1" and RescaleIntercept "0" present: no rescale (a window
    # wide enough that the rescaled and the plain codes give different rows)
    pw = r2p.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 500.0, 3000.0)
    synth.write_dicom(str(tmp_path / "r10.dcm"), px, pw, rescale=True, slope_str=b"1", intercept_str=b"0")
    p_norescale = r2p.DicomParams.make(16, 12, 11, 0, 0, 1.0, 0.0, 1, 500.0, 3000.0)
    got = encode("dicom", str(tmp_path / "r10.dcm"), "r10.png")
    assert np.array_equal(got, expect(p_norescale)) and not np.array_equal(got, expect(pw))
