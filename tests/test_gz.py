"""gz_decode (the reference's gz::decode member loop, src/gz.rs:154-264) on the GPU: multi-member
files with every header field, trailing data, and each failure class, against Python's gzip."""
import gzip
import struct
import zlib

import pytest

from gpuzip_amd import datagen
from gpuzip_amd import gz
from test_gz_header import member

pytestmark = pytest.mark.gpu


def test_members_and_headers(gpu):
    parts = [datagen.text(3 << 20, seed=1).tobytes(), datagen.mixed(2 << 20, seed=2).tobytes(), b"", b"tail " * 999]
    bc = b"BC" + struct.pack("<HH", 2, 0)
    heads = [dict(), dict(flags=8, name=b"a.txt"), dict(flags=16 | 2, comment=b"c", hcrc=True),
             dict(flags=4, extra=bc)]
    data = b"".join(b"".join(member(p, **h)) for p, h in zip(parts, heads))
    out, rep = gz.decode(gpu, data)
    assert out == b"".join(parts) == gzip.decompress(data)
    assert rep.members == 4 and rep.out == len(out) and rep.packed == len(data)
    assert rep.info == b"" and rep.note == b""


def test_trailing_data(gpu):
    p = datagen.text(1 << 20, seed=3).tobytes()
    one = gzip.compress(p, mtime=0)
    out, rep = gz.decode(gpu, one + bytes(100))
    assert out == p and rep.info.decode() == "100 zero bytes after the last member"
    out, rep = gz.decode(gpu, one + b"garbage")
    assert out == p and rep.note.decode() == "7 bytes after the last member are not gzip data; ignored"


def test_failures(gpu):
    p = datagen.text(1 << 20, seed=4).tobytes()
    one = bytearray(gzip.compress(p, mtime=0))
    with pytest.raises(gz.Corrupt, match="no gzip member"):
        gz.decode(gpu, b"")
    with pytest.raises(gz.Corrupt, match="not a gzip member"):
        gz.decode(gpu, b"plain text, not gzip")
    bad = bytearray(one)
    bad[-8] ^= 1  # the trailer's CRC
    with pytest.raises(gz.Retry, match="CRC check"):
        gz.decode(gpu, bytes(bad))
    bad = bytearray(one)
    bad[-1] ^= 1  # its size
    with pytest.raises(gz.Retry, match="size check"):
        gz.decode(gpu, bytes(bad))
    with pytest.raises(gz.Corrupt):
        gz.decode(gpu, bytes(one[: len(one) // 2]))
    with pytest.raises(gz.Corrupt, match="trailer"):
        gz.decode(gpu, bytes(one[:-4]))
    with pytest.raises(gz.Stopped):
        gz.decode(gpu, bytes(one), sink=lambda b: False)
    # the context is still usable
    out, rep = gz.decode(gpu, bytes(one))
    assert out == p and rep.members == 1


def test_fault_hook_is_retry_and_context_survives(monkeypatch):
    """GI_FAULT_AFTER_SEGMENTS (read when a context is made) fails a call as it begins segment
    N + 1 with GI_ERR_GPU: gz_decode maps it to Retry (the caller's CPU fallback, src/gz.rs:229),
    and the same context then decodes a stream of N segments or fewer bit-exact"""
    from gpuzip_amd.gpuinflate import Gpu, GpuError
    monkeypatch.setenv("GI_FAULT_AFTER_SEGMENTS", "1")
    big = datagen.text(40 << 20, seed=5).tobytes()
    small = datagen.text(1 << 20, seed=6).tobytes()
    with Gpu(segment_out_mb=16) as g:
        with pytest.raises(gz.Retry, match="injected fault after 1 segments"):
            gz.decode(g, gzip.compress(big, mtime=0))
        with pytest.raises(GpuError):
            g.inflate(zlib.compress(big)[2:-4], len(big))
        out, rep = gz.decode(g, gzip.compress(small, mtime=0))
        assert out == small and rep.members == 1
    monkeypatch.delenv("GI_FAULT_AFTER_SEGMENTS")
    with Gpu(segment_out_mb=16) as g:
        out, rep = gz.decode(g, gzip.compress(big, mtime=0))
        assert out == big and rep.stats.segments >= 3
