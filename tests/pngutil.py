"""PNG container parsing for the tests (independent of the encoder)."""
import struct
import zlib


def chunks(png: bytes):
    assert png[:8] == b"\x89PNG\r\n\x1a\n", "bad signature"
    pos, out = 8, []
    while pos < len(png):
        (n,) = struct.unpack(">I", png[pos:pos + 4])
        typ = png[pos + 4:pos + 8]
        data = png[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", png[pos + 8 + n:pos + 12 + n])
        out.append((typ, data, crc))
        pos += 12 + n
    assert pos == len(png), "trailing bytes"
    return out


def check_and_inflate(png: bytes):
    """Validates every chunk CRC with zlib.crc32 and returns (ihdr, zlib stream, inflated)."""
    cs = chunks(png)
    for typ, data, crc in cs:
        assert zlib.crc32(typ + data) == crc, f"bad CRC in {typ!r}"
    assert cs[0][0] == b"IHDR" and cs[-1][0] == b"IEND" and cs[-1][1] == b""
    w, h, depth, ctype, comp, filt, inter = struct.unpack(">IIBBBBB", cs[0][1])
    z = b"".join(d for t, d, _ in cs if t == b"IDAT")
    assert (z[0] * 256 + z[1]) % 31 == 0 and z[0] == 0x78
    d = zlib.decompressobj()
    raw = d.decompress(z) + d.flush()
    assert d.eof and d.unused_data == b"", "stream not terminated exactly"
    assert struct.unpack(">I", z[-4:])[0] == zlib.adler32(raw)
    return dict(width=w, height=h, depth=depth, ctype=ctype), z, raw
