"""Pins the oracle (oracle/inflate_oracle.c) before it is trusted as the checker:
- CRC-32 known answer and zlib.crc32;
- whole-stream decode against system zlib (bytes and consumed input) on every stream kind;
- against the reference's own decoder compiled from its sources (oracle/_ref), on fuzzed
  streams: status, bytes and in_used, and Decoder::run result fields with landing targets;
- the block-start predicates against the reference's, bit by bit, and against block starts
  that zlib's full flushes put at known positions (real headers always pass).
"""
import ctypes
import random
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from oracle_lib import OrResult, OrStart, oracle, oracle_inflate, ref, ref_inflate
import reference_corpus as refcorp

u64 = ctypes.c_uint64


def test_crc32_known_answer():
    o = oracle()
    assert o.or_crc32(0, b"123456789", 9) == 0xCBF43926
    assert o.or_crc32(0, b"", 0) == 0
    rnd = random.Random(1).randbytes(100_000)
    assert o.or_crc32(0, rnd, len(rnd)) == zlib.crc32(rnd)
    # running value: crc(a + b) = crc(b, crc(a))
    a, b = rnd[:777], rnd[777:]
    assert o.or_crc32(o.or_crc32(0, a, len(a)), b, len(b)) == zlib.crc32(rnd)
    assert datagen.crc32(np.frombuffer(rnd, np.uint8)) == zlib.crc32(rnd)


def test_oracle_matches_zlib_on_every_kind(corpus):
    for name, (stream, original) in corpus.items():
        tail = b"\x00trailer-and-more"
        rc, out, used, big = oracle_inflate(stream + tail)
        assert rc == 0 and not big, name
        assert out == original, name
        d = zlib.decompressobj(-15)
        assert d.decompress(stream + tail) == original
        assert used == len(stream + tail) - len(d.unused_data), name


def test_oracle_matches_zlib_on_levels_and_strategies():
    data = datagen.mixed(3 << 20, seed=11, run_max=1 << 20)
    for level in (0, 1, 2, 4, 6, 9):
        for strategy in (0, 1, 2, 3, 4):  # default, filtered, huffman only, RLE, fixed
            s = bytes(datagen.deflate(data, level, strategy))
            rc, out, used, _ = oracle_inflate(s)
            assert rc == 0 and out == data.tobytes() and used == len(s), (level, strategy)


def test_oracle_error_classes_known_answers():
    # empty stored block, final: decodes to nothing; in_used = 5
    assert oracle_inflate(bytes([1, 0, 0, 0xFF, 0xFF]))[:3] == (0, b"", 5)
    # BTYPE 3 is invalid data
    assert oracle_inflate(bytes([0x07]))[0] == 1
    # a stored block whose LEN/NLEN disagree
    assert oracle_inflate(bytes([1, 5, 0, 0, 0]))[0] == 1
    # a stored block whose bytes run past the input is truncated
    assert oracle_inflate(bytes([1, 10, 0, 0xF5, 0xFF, 1, 2, 3]))[0] == 2
    # so is a stream that ends without its last block
    s = bytes(datagen.deflate(datagen.text(1 << 20, seed=2), 6))
    assert oracle_inflate(s[: len(s) // 2])[0] == 2
    # a match reaching before the start: fixed block, length 3 distance 1 as the first symbol
    # (length code 257 = 0000001, 7 bits; distance code 0 = 00000) -> data error
    bits = [1, 1, 0] + [0, 0, 0, 0, 0, 0, 1] + [0, 0, 0, 0, 0] + [0, 0, 0, 0, 0, 0, 0]
    v = 0
    for i, b in enumerate(bits):
        v |= b << i
    assert oracle_inflate(v.to_bytes(4, "little"))[0] == 1


@pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (reference sources absent)")
def test_oracle_agrees_with_reference_decoder_on_fuzzed_streams(corpus):
    rng = refcorp.Splitmix(1)
    n = 0
    for name in ("f-l1", "f-l9", "f-fixed", "f-stored", "f-mixed", "f-huffman", "f-tiny"):
        stream, _ = corpus[name]
        stream = stream[: 96 << 10] if name != "f-tiny" else stream
        for _ in range(40):
            m = refcorp.mutate(stream, rng)
            a = oracle_inflate(m, 8 << 20)
            b = ref_inflate(m, 8 << 20)
            assert a[3] == b[3]
            if not a[3]:
                assert a[0] == b[0], name
                if a[0] == 0:
                    assert a[1] == b[1] and a[2] == b[2], name
            n += 1
    assert n == 280


@pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (reference sources absent)")
def test_predicates_agree_with_reference_bit_by_bit():
    o, r = oracle(), ref()
    rng = random.Random(5)
    for _ in range(20000):
        v1, v2 = rng.getrandbits(64), rng.getrandbits(64)
        # bias toward passing the field tests so the Kraft part is exercised
        v1 = (v1 & ~0b110) | 0b100
        assert o.or_header_prefilter(v1) == r.ref_header_prefilter(v1)
        assert o.or_header_prefilter128(v1, v2) == r.ref_header_prefilter128(v1, v2)
    stream = bytes(datagen.deflate(datagen.text(256 << 10, seed=3), 9))[:40000]
    hits = 0
    for b in range(0, len(stream) * 8, 3):
        x = o.or_plausible_dynamic_header(stream, len(stream), b)
        assert x == r.ref_plausible_dynamic_header(stream, len(stream), b), b
        hits += x
    assert hits > 0


def test_real_headers_pass_the_predicates():
    """full flushes put block starts at known byte positions; each following dynamic block's
    header passes every predicate (zlib accepts it, so plausible_dynamic_header must)"""
    o = oracle()
    t = datagen.text(2 << 20, seed=9).tobytes()
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    out, starts = b"", []
    for i in range(0, len(t), 64 << 10):
        out += c.compress(t[i:i + (64 << 10)])
        out += c.flush(zlib.Z_FULL_FLUSH)
        starts.append(len(out) * 8)
    out += c.flush()
    for b in starts[:-1]:
        v1, v2 = o.or_bits64(out, len(out), b), o.or_bits64(out, len(out), b + 64)
        assert (v1 >> 1) & 3 == 2  # text at level 6 is a dynamic block
        assert o.or_header_prefilter(v1) and o.or_header_prefilter128(v1, v2)
        assert o.or_plausible_dynamic_header(out, len(out), b)
    found = np.empty(4096, np.uint64)
    n = o.or_find_headers(out, len(out), 0, len(out) * 8, found.ctypes.data, 4096)
    assert set(starts[:-1]) <= set(int(x) for x in found[:n])


@pytest.mark.skipif(ref() is None, reason="oracle/_ref not built (reference sources absent)")
def test_run_agrees_with_reference_with_landing_targets(corpus):
    """Decoder::run from every block start, landing on the later ones: every result field"""
    o, r = oracle(), ref()
    for name in ("f-l6", "f-fixed", "f-mixed"):
        stream = corpus[name][0][: 256 << 10]
        found = np.empty(1 << 16, np.uint64)
        n = o.or_find_headers(stream, len(stream), 1, len(stream) * 8, found.ctypes.data, 1 << 16)
        starts = [OrStart(0, 0, 0, -1)] + [OrStart(int(b), int(b), 0, -1) for b in found[:n]]
        targets = (OrStart * len(starts))(*starts)
        for k in range(0, len(starts), max(1, len(starts) // 12)):
            tg = ctypes.cast(ctypes.byref(targets, (k + 1) * ctypes.sizeof(OrStart)), ctypes.POINTER(OrStart))
            nt = len(starts) - k - 1
            for cap in (1 << 20, 5000):
                out_a = np.zeros(cap, np.uint16)
                out_b = np.zeros(cap, np.uint16)
                ra = o.or_run(stream, len(stream), starts[k], tg, nt, out_a.ctypes.data, cap, int(k == 0), 2**64 - 1)
                rb = OrResult()
                r.ref_run(stream, len(stream), ctypes.byref(starts[k]), tg, nt, out_b.ctypes.data, cap, int(k == 0),
                          2**64 - 1, ctypes.byref(rb))
                for f, _ in OrResult._fields_:
                    if f == "resume":
                        if ra.status not in (2, 4):  # set only when the decoder stopped for room or at the stop bit
                            continue
                        for g, _ in OrStart._fields_:
                            assert getattr(ra.resume, g) == getattr(rb.resume, g), (name, k, f, g)
                    else:
                        assert getattr(ra, f) == getattr(rb, f), (name, k, cap, f)
                assert (out_a[: ra.out_len] == out_b[: rb.out_len]).all()
