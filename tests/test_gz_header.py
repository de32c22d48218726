"""The gzip member header parser (gz_header, the reference's parse_header, src/gz.rs:62-110),
pinned by RFC 1952's layout and by zlib: for every valid header, raw Deflate decoding starting at
the reported length gives the member's data back. No GPU needed."""
import struct
import zlib

import pytest

from gpuzip_amd import gz


def member(payload, flags=0, name=b"", comment=b"", extra=b"", hcrc=False, method=8, level=6):
    h = bytearray(b"\x1f\x8b" + bytes([method, flags]) + struct.pack("<I", 0) + b"\x00\x03")
    if flags & 4:
        h += struct.pack("<H", len(extra)) + extra
    if flags & 8:
        h += name + b"\x00"
    if flags & 16:
        h += comment + b"\x00"
    if flags & 2:
        h += struct.pack("<H", zlib.crc32(bytes(h)) & 0xFFFF if hcrc else (zlib.crc32(bytes(h)) + 1) & 0xFFFF)
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    body = c.compress(payload) + c.flush()
    return bytes(h), body + struct.pack("<II", zlib.crc32(payload), len(payload) & 0xFFFFFFFF)


def check_valid(h, rest, payload, want_member=0):
    n, m = gz.header(h + rest)
    assert n == len(h) and m == want_member
    d = zlib.decompressobj(-15)
    assert d.decompress((h + rest)[n:]) == payload


def test_valid_headers():
    payload = b"hello gzip " * 100
    h, r = member(payload)
    check_valid(h, r, payload)
    h, r = member(payload, flags=8, name=b"file.txt")
    check_valid(h, r, payload)
    h, r = member(payload, flags=16, comment=b"a comment")
    check_valid(h, r, payload)
    h, r = member(payload, flags=8 | 16 | 2, name=b"n", comment=b"c", hcrc=True)
    check_valid(h, r, payload)
    # BGZF: the BC subfield holds the member's size less one (here after another subfield)
    bc = b"BC" + struct.pack("<HH", 2, 1234 - 1)
    h, r = member(payload, flags=4, extra=b"XY" + struct.pack("<H", 3) + b"abc" + bc)
    check_valid(h, r, payload, want_member=1234)
    h, r = member(payload, flags=4, extra=b"XY" + struct.pack("<H", 0))
    check_valid(h, r, payload)


@pytest.mark.parametrize("case,msg", [
    ("short", "not a gzip member"),
    ("magic", "not a gzip member"),
    ("method", "compression method 7 is not Deflate"),
    ("reserved", "reserved header flags are set"),
    ("hcrc", "header CRC mismatch"),
])
def test_bad_headers(case, msg):
    payload = b"x" * 50
    if case == "short":
        data = b"\x1f\x8b\x08\x00\x00"
    elif case == "magic":
        h, r = member(payload)
        data = b"\x1f\x8c" + (h + r)[2:]
    elif case == "method":
        h, r = member(payload, method=7)
        data = h + r
    elif case == "reserved":
        h, r = member(payload, flags=0x20)
        data = h + r
    else:
        h, r = member(payload, flags=2, hcrc=False)
        data = h + r
    with pytest.raises(gz.Corrupt, match=msg):
        gz.header(data)


def test_every_truncation_of_a_full_header():
    h, _ = member(b"y", flags=4 | 8 | 16 | 2, name=b"name", comment=b"comment", extra=b"AB" + struct.pack("<H", 2) + b"zz",
                  hcrc=True)
    for k in range(len(h)):
        with pytest.raises(gz.Corrupt, match="not a gzip member" if k < 10 else "the file ends inside a member header"):
            gz.header(h[:k])
    assert gz.header(h)[0] == len(h)
