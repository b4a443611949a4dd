"""The C-ABI library loads and exports every function include/*.h declares.
CPU only: no compute calls."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from conftest import PKG_DIR, REPO


def declared_c_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        txt = open(h).read()
        if 'extern "C"' not in txt and "#ifdef __cplusplus" not in txt:
            continue
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"\b(r2p_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_all_declared(r2p):
    lib = r2p.lib()
    names = declared_c_functions()
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", r2p.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (r2p_\w+)", out))
    assert names <= exported


def test_geometry_helpers_without_device(r2p):
    p = r2p.DicomParams.make(16, 12, 11, 0, 1, 1.0, -1024.0, 1, 40.0, 400.0)
    img = r2p.Image.make(512, 512, 1, 16, r2p.PIX_DICOM8, p)
    assert r2p.out_bits(img) == 8
    assert r2p.row_bytes(img) == 512
    assert r2p.filtered_bytes(img) == 512 * 513 == 262656
    img16 = r2p.Image.make(3072, 3072, 1, 16, r2p.PIX_DICOM, p)
    assert r2p.filtered_bytes(img16) == 3072 * 6145 == 18877440
    rgb = r2p.Image.make(6000, 4000, 3, 8, r2p.PIX_READY)
    assert r2p.filtered_bytes(rgb) == 72004000
    assert r2p.input_bytes(rgb) == 72000000
    assert r2p.png_bound(rgb) > r2p.filtered_bytes(rgb) * 9 // 8
    bad = r2p.Image.make(0, 5, 1, 8, 0)
    assert r2p.filtered_bytes(bad) == 0
    assert r2p.out_bits(r2p.Image.make(5, 5, 1, 8, r2p.PIX_SWAP16)) == -1


def test_options_default(r2p):
    o = r2p.options()
    assert (o.segment_bytes, o.lz77, o.idat_chunk_bytes, o.zlib_level) == (65536, 1, 65536, 3)
    assert (o.lane_bytes, o.wg_threads, o.slice_bytes) == (16, 256, 0)    # slices: by image size


def test_host_combines_match_zlib(r2p):
    import zlib
    a, b = b"hello world" * 1000, b"segment two" * 777
    assert r2p.lib().r2p_adler32_combine(zlib.adler32(a), zlib.adler32(b), len(b)) == zlib.adler32(a + b)
    assert r2p.lib().r2p_crc32_combine(zlib.crc32(a), zlib.crc32(b), len(b)) == zlib.crc32(a + b)


def test_no_device_errors_cleanly(r2p):
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("device present")
    with pytest.raises(r2p.R2PError) as ei:
        r2p.Encoder(1000)
    assert ei.value.code == r2p.R2P_ERR_NODEV
