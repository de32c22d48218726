"""Parity of the MI355X decoder (libgpuinflate.so, through the C interface) with the oracle:
bit-exact bytes, the same in_used (where the gzip trailer is found), the same error class, and
the GPU's CRC-32; the reference's own checks (gpuinflate/README.md:53-62): every stream kind,
thousands of segment boundaries, fuzzed streams."""
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from gpuzip_amd import gpuinflate as gi
from oracle_lib import oracle_inflate
import reference_corpus as refcorp

pytestmark = pytest.mark.gpu


def check(g, stream, original, hint=0):
    tail = b"\x1f\x8b-trailer"
    out, st = g.inflate(stream + tail, hint)
    assert len(out) == len(original)
    assert out == original
    assert st.in_used == len(stream)
    assert st.out_size == len(original)
    assert g.last_crc32() == zlib.crc32(original)
    return st


def test_every_kind(gpu, corpus):
    for name, (stream, original) in corpus.items():
        print("kind", name, len(stream), flush=True)
        check(gpu, stream, original)


@pytest.mark.parametrize("seg_in_kb", [16, 32])
def test_tiny_segments(corpus, seg_in_kb):
    """thousands of segment boundaries: windows, headers before the segment, BFINAL carried"""
    with gi.Gpu(segment_out_mb=8, segment_in_kb=seg_in_kb) as g:
        for name in ("f-l6", "f-fixed", "f-mixed", "f-huffman", "f-stored", "f-tiny"):
            stream, original = corpus[name]
            print("tiny", seg_in_kb, name, flush=True)
            st = check(g, stream, original)
            if len(stream) > (seg_in_kb << 12):
                assert st.segments > 2


def test_log_text_64mib_level6(gpu):
    """the CPU case's content and size: 64 MiB of log text, zlib level 6"""
    t = datagen.text(64 << 20)
    s = bytes(datagen.deflate(t, 6))
    out, st = gpu.inflate(s, len(t))
    assert out == t.tobytes() and st.in_used == len(s)
    assert gpu.last_crc32() == datagen.crc32(t)


def test_false_stored_candidate(gpu):
    """fixed-code data that reads as a stored block whose end is a real stored block's start: the
    decode goes on past the false candidate"""
    s, d, _ = datagen.false_stored_stream()
    check(gpu, s, d)
    with gi.Gpu(segment_out_mb=8, segment_in_kb=64) as g:
        check(g, s, d)


def test_mixed_levels(gpu):
    data = datagen.mixed(24 << 20, seed=77, run_max=2 << 20)
    for level in (1, 6, 9):
        s = bytes(datagen.deflate(data, level, 0, chunk=4 << 20))
        print("level", level, flush=True)
        check(gpu, s, data.tobytes())


@pytest.mark.parametrize("seg_in_kb", [0, 16])
def test_fuzz_agrees_with_oracle(gpu, corpus, seg_in_kb):
    """gi_test fuzz (gi_test.cpp:192-251) at fuzz_all.ps1's scale: every corpus kind, 300 damaged
    copies each, with the default segments and with 16 KiB input segments (many segment
    boundaries): 4,800 decodes in all. Each gives the oracle's error class, or its bytes and
    in_used. Streams are cut to their first 256 KiB so that the oracle keeps up."""
    ctx = gpu if seg_in_kb == 0 else gi.Gpu(device=0, segment_in_kb=seg_in_kb)
    counts = {"ok": 0, "data": 0, "truncated": 0, "big": 0}
    try:
        for name in ("f-l1", "f-l6", "f-l9", "f-fixed", "f-stored", "f-mixed", "f-huffman", "f-tiny"):
            print("kind", name, "segments", seg_in_kb or "default", flush=True)
            rng = refcorp.Splitmix(1)
            stream = corpus[name][0]
            stream = stream[: 256 << 10] if len(stream) > (256 << 10) else stream
            for it in range(300):
                m = refcorp.mutate(stream, rng)
                want = oracle_inflate(m, 32 << 20)
                if want[3]:
                    counts["big"] += 1
                    continue
                try:
                    out, st = ctx.inflate(m)
                    got = (0, out, st.in_used)
                except gi.DataError:
                    got = (1, b"", 0)
                except gi.TruncatedError:
                    got = (2, b"", 0)
                assert got[0] == want[0], (name, it, ctx.last_error())
                counts[("ok", "data", "truncated")[want[0]]] += 1
                if want[0] == 0:
                    assert got[1] == want[1] and got[2] == want[2], (name, it)
    finally:
        if ctx is not gpu:
            ctx.close()
    print("fuzz", seg_in_kb or "default", counts, flush=True)
    assert counts["ok"] > 20 and counts["data"] > 100 and counts["truncated"] > 100


def test_errors_and_abort(gpu, corpus):
    with pytest.raises(gi.TruncatedError):
        gpu.inflate(b"")
    with pytest.raises(gi.DataError):
        gpu.inflate(bytes([0x07, 0, 0, 0]))
    stream, original = corpus["f-l6"]
    with pytest.raises(gi.TruncatedError):
        gpu.inflate(stream[: len(stream) // 2])
    calls = []

    def stop_after_one(b):
        calls.append(len(b))
        return False

    with pytest.raises(gi.AbortedError):
        gpu.inflate(stream, sink=stop_after_one)
    assert len(calls) == 1
    # the context stays usable after errors
    check(gpu, stream, original)


def test_device_resident(gpu):
    """gi_inflate_device: input and output in device memory; the same bytes and CRC-32"""
    from gpuzip_amd.hipmem import DeviceBuffer
    data = datagen.mixed(32 << 20, seed=5)
    s = datagen.deflate(data, 6, 0, chunk=8 << 20)
    d_in = DeviceBuffer.from_host(s)
    d_out = DeviceBuffer(data.size)
    n, crc, st = gpu.inflate_device(d_in.addr, s.size, d_out.addr, data.size)
    assert n == data.size and st.in_used == s.size
    assert crc == datagen.crc32(data)
    assert np.array_equal(d_out.to_host(), data)
    with pytest.raises(gi.ArgError):
        gpu.inflate_device(d_in.addr, s.size, d_out.addr, data.size - 1)
    # a second stream through the same context, output at an odd offset
    t = datagen.text(5 << 20, seed=8)
    s2 = datagen.deflate(t, 9)
    d_in2 = DeviceBuffer.from_host(s2)
    n, crc, st = gpu.inflate_device(d_in2.addr, s2.size, d_out.addr + 3, data.size - 3)
    assert n == t.size and crc == datagen.crc32(t)
    assert np.array_equal(d_out.to_host(t.size + 3)[3:], t)


def test_finder_graph_replays(monkeypatch):
    """the finder's launches and the resolve's run as one HIP graph each per segment, updated in
    place from segment to segment and replayed with the decode's own launches and host waits
    between: over many small segments and two calls on one context, the output equals zlib's and
    that of the kernels launched one by one (GZ_NO_GRAPH)"""
    data = datagen.mixed(24 << 20, seed=31, run_max=2 << 20).tobytes()
    s = bytes(datagen.deflate(np.frombuffer(data, np.uint8), 6, 0, chunk=4 << 20))
    outs = {}
    for mode in ("graph", "direct"):
        if mode == "direct":
            monkeypatch.setenv("GZ_NO_GRAPH", "1")
        with gi.Gpu(segment_out_mb=8, segment_in_kb=512) as g:
            for call in range(2):
                out, st = g.inflate(s)
                assert out == data and st.in_used == len(s) and g.last_crc32() == zlib.crc32(data), (mode, call)
                assert st.segments > 4
            names = set(g.kernel_times())
        outs[mode] = names
        monkeypatch.delenv("GZ_NO_GRAPH", raising=False)
    assert "find_graph" in outs["graph"] and "k_find" not in outs["graph"]
    assert "resolve_graph" in outs["graph"] and "k_tile_a" not in outs["graph"]
    assert "k_find" in outs["direct"] and "find_graph" not in outs["direct"]
    assert "k_tile_a" in outs["direct"] and "resolve_graph" not in outs["direct"]


@pytest.mark.timeout(900)
def test_one_deflate_call_1gib(gpu):
    """a 1 GiB stream of the 4 GiB workload's content made by ONE zlib deflate() call at level 6
    (no chunks, no sync-flush points): bit-exact, in_used and CRC-32 equal zlib's"""
    data = datagen.mixed(1 << 30, seed=91)
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    print("compressing 1 GiB in one deflate() call", flush=True)
    s = c.compress(data.tobytes()) + c.flush()
    print("compressed to", len(s), flush=True)
    from gpuzip_amd.hipmem import DeviceBuffer
    d_in = DeviceBuffer.from_host(np.frombuffer(s, np.uint8))
    d_out = DeviceBuffer(data.size)
    n, crc, st = gpu.inflate_device(d_in.addr, len(s), d_out.addr, data.size, 0)
    assert n == data.size and st.in_used == len(s)
    assert crc == datagen.crc32(data)
    assert np.array_equal(d_out.to_host(), data)


@pytest.mark.timeout(900)
def test_output_over_4gib_size_wraps():
    """4.5 GiB of output in 200 MiB segments (over 20 boundaries; 64-bit positions throughout):
    the gzip size field wraps to a non-zero value (the reference's make_big_gz.py check,
    gpuinflate/README.md:62); every byte and the CRC-32 equal the original's"""
    from gpuzip_amd.hipmem import DeviceBuffer
    U = (4 << 30) + (512 << 20)
    data = datagen.mixed(U, seed=92)
    s = datagen.deflate(data, 4, 0, chunk=64 << 20)
    print("4.5 GiB compressed to", s.size, flush=True)
    want_crc = datagen.crc32(data)
    isize = U & 0xFFFFFFFF
    assert isize == 512 << 20
    d_in = DeviceBuffer.from_host(s)
    d_out = DeviceBuffer(U)
    with gi.Gpu(segment_out_mb=200) as g:
        n, crc, st = g.inflate_device(d_in.addr, s.size, d_out.addr, U, isize)
    assert st.segments >= 20, st.segments
    assert n == U and (n & 0xFFFFFFFF) == isize and crc == want_crc and st.in_used == s.size
    assert np.array_equal(d_out.to_host(), data)
