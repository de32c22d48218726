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


def test_many_streams_against_zlib():
    """gi_inflate_many: streams of mixed sizes and levels from one host buffer, among them a damaged
    one, one without enough output room, and one too large to share a batch segment (decoded
    alone: the context has 64 MiB segments, a quarter of which is the most a shared stream gets);
    each stream's results are those of gi_inflate on it alone"""
    from gpuzip_amd import gpuinflate as gi
    with gi.Gpu(device=0, segment_out_mb=64) as gpu:
        many_streams(gpu, gi)


def many_streams(gpu, gi):
    import random
    rng = random.Random(5)
    raw, spans, want, at = [], [], [], 0
    for k in range(120):
        n = rng.choice([0, 1, 100, 5000, 70000, 400000])
        d = datagen.mixed(n, seed=k, run_max=1 << 16).tobytes() if n > 64 else bytes(n)
        s = bytes(datagen.deflate(d, rng.choice([1, 6, 9]))) if n else zlib.compress(b"", 6)[2:-4]
        raw.append(s)
        want.append(d)
        spans.append([at, len(s), len(d)])
        at += len(s) + rng.randrange(0, 40)
        raw.append(bytes(at - sum(len(x) for x in raw)))
    big = datagen.mixed(40 << 20, seed=99).tobytes()
    bs = bytes(datagen.deflate(big, 6, chunk=8 << 20))
    spans.append([at, len(bs), len(big)])
    raw.append(bs)
    want.append(big)
    data = bytearray(b"".join(raw))
    assert len(data) == at + len(bs)
    bad = next(i for i in range(17, 120) if spans[i][1] > 100)
    tight = next(i for i in range(bad + 1, 120) if spans[i][2] > 0)
    data[spans[bad][0] + spans[bad][1] // 2] ^= 0xFF
    spans[tight][2] -= 1
    print("streams", len(spans), "bytes", len(data), flush=True)
    res, outs, rc = gpu.inflate_many(bytes(data), spans)
    assert rc != gi.GI_OK
    for i, (r, o) in enumerate(zip(res, outs)):
        off, size, room = spans[i]
        try:
            alone, st = gpu.inflate(bytes(data[off:off + size]), room)
            alone_rc = gi.GI_OK
        except gi.Error as e:
            alone, alone_rc = None, e.code
        if i == tight:
            assert r.status == gi.GI_ERR_ARG, i
            continue
        assert r.status == alone_rc, (i, r.status, alone_rc)
        if r.status == gi.GI_OK:
            assert o == want[i] == alone and r.crc32 == zlib.crc32(want[i]) and r.in_used == st.in_used, i
    assert res[bad].status in (gi.GI_ERR_DATA, gi.GI_ERR_TRUNCATED)


def test_bgzf_file_of_301_members(gpu):
    """a bgzip file (300 members of 64 KB, each one's size in its BC subfield, and the empty member
    at the end: the reference's test, tests/test_gpuzip.py:1008-1040) decodes whole in a few
    gi_inflate_many segments; against one gi_inflate a member it takes a fraction of the time"""
    import random
    import time
    from test_gzhost_cpu import bgzf_blocks, bgzf_member
    blocks = bgzf_blocks(300)
    data = b"".join(bgzf_member(b) for b in blocks) + bgzf_member(b"")
    gz.decode(gpu, b"".join(bgzf_member(b) for b in blocks[:3]))  # warm-up
    t = time.perf_counter()
    out, rep = gz.decode(gpu, data)
    batched = time.perf_counter() - t
    assert out == b"".join(blocks) and rep.members == 301
    t = time.perf_counter()
    one = [gpu.inflate(bgzf_member(b)[18:-8], len(b))[0] for b in blocks[:30]]
    per_member = (time.perf_counter() - t) / 30
    assert b"".join(one) == b"".join(blocks[:30])
    print(f"301 members: {batched * 1e3:.1f} ms batched, {per_member * 1e3:.2f} ms a member alone", flush=True)
    assert batched < 301 * per_member / 3
