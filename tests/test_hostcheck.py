"""The decoding core and planner that the GPU library runs (csrc/deflate_core.h, csrc/plan.h),
built for the host with a CPU backend (libgzhostcheck.so), checked against the oracle: the
same segment loop, stitching and decoding logic as the kernels, without a GPU."""
import ctypes
import random
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from oracle_lib import hc_inflate, hc_inflate_batch, hostcheck, oracle, oracle_inflate
import reference_corpus as refcorp


def test_struct_layouts():
    h = hostcheck()
    assert h.hc_sizeof(0) == 24 and h.hc_sizeof(1) == 72 and h.hc_sizeof(2) % 16 == 0


@pytest.mark.parametrize("seg", [(256 << 20, 0), (8 << 20, 32), (8 << 20, 16)])
def test_every_kind_bit_exact(corpus, seg):
    for name, (stream, original) in corpus.items():
        tail = b"TRAILER!"
        rc, out, used, segs, rounds, err = hc_inflate(stream + tail, 0, seg[0], seg[1])
        assert rc == 0, (name, err)
        assert out == original, name
        assert used == len(stream), name


def test_mixed_levels_and_hint():
    data = datagen.mixed(6 << 20, seed=21, run_max=2 << 20)
    for level in (1, 6, 9):
        s = bytes(datagen.deflate(data, level, 0, chunk=1 << 20))
        for hint in (0, data.size):
            rc, out, used, *_ = hc_inflate(s, hint)
            assert rc == 0 and out == data.tobytes() and used == len(s), (level, hint)


def test_false_stored_candidate():
    """a fixed block holding four bytes that read as a stored block's LEN/NLEN, whose end is where
    a real stored block starts (so the planner takes the candidate for a stored block): the group
    over that block stops where the false LEN lies and goes on in a continuation"""
    s, d, at = datagen.false_stored_stream()
    n = s[at] | s[at + 1] << 8
    assert n == (~(s[at + 2] | s[at + 3] << 8) & 0xFFFF)
    q = at + 4 + n
    assert s[q + 1] | s[q + 2] << 8 == (~(s[q + 3] | s[q + 4] << 8) & 0xFFFF)
    assert zlib.decompress(s, -15) == d
    for seg in ((256 << 20, 0), (8 << 20, 64)):
        rc, out, used, *_, err = hc_inflate(s, 0, seg[0], seg[1])
        assert rc == 0 and out == d and used == len(s), (seg, err)


def test_batch_matches_alone(corpus):
    """many streams as one batch segment (packed in one input, one plan, one resolve): every
    stream's bytes, in_used and CRC-32 as decoded alone, damaged and truncated ones included,
    and the output space covered exactly (the host build checks it)"""
    from gpuzip_amd import workloads
    streams = [s for s, _ in corpus.values()]
    pool = workloads.content_pool(8 << 20)
    sizes = np.minimum(workloads.batch_sizes(40, seed=3), 2 << 20)
    streams += [s for s, _, _ in workloads.make_streams(sizes, pool)]
    streams.append(zlib.compressobj(6, zlib.DEFLATED, -15).flush())  # empty output
    l6, l1 = corpus["f-l6"][0], bytearray(corpus["f-l1"][0])
    streams.append(l6[: len(l6) // 2])  # truncated
    l1[100] ^= 0xFF
    streams.append(bytes(l1))  # damaged
    streams.append(datagen.false_stored_stream()[0])
    caps = [64 << 20] * len(streams)
    res = hc_inflate_batch(streams, caps)
    for k, (s, (rc, out, used, crc)) in enumerate(zip(streams, res)):
        arc, aout, aused, *_ = hc_inflate(s, 0, 256 << 20, 0, 64 << 20)
        assert rc == arc, (k, rc, arc)
        if rc == 0:
            assert out == aout and used == aused and crc == zlib.crc32(aout), k
    assert [r[0] for r in res[-4:]] == [0, 2, 1, 0]


def test_finder_equals_brute_force():
    """the word-sliced prefilter plus zlib's checks finds exactly the positions a bit-by-bit
    application of the predicates finds"""
    o, h = oracle(), hostcheck()
    for seed, level in ((1, 6), (2, 9), (3, 1)):
        s = bytes(datagen.deflate(datagen.mixed(1 << 20, seed=seed, run_max=200 << 10), level))[:300_000]
        nb = len(s) * 8
        for lo, hi in ((0, nb), (12345, nb - 777)):
            a = np.empty(1 << 16, np.uint64)
            b = np.empty(1 << 16, np.uint64)
            na = h.hc_find(s, len(s), lo, hi, a.ctypes.data, 1 << 16)
            nb_ = o.or_find_headers(s, len(s), lo + 1, hi, b.ctypes.data, 1 << 16)
            assert sorted(int(x) for x in a[:na]) == [int(x) for x in b[:nb_]]


def test_prefilter_equals_predicate_on_random_bits():
    o, h = oracle(), hostcheck()
    buf = random.Random(4).randbytes(20000)
    out = np.empty(1 << 16, np.uint64)
    n = h.hc_prefilter(buf, len(buf), 0, len(buf) * 8, out.ctypes.data, 1 << 16)
    got = set(int(x) for x in out[:n])
    want = set()
    for b in range(1, len(buf) * 8):
        v1, v2 = o.or_bits64(buf, len(buf), b), o.or_bits64(buf, len(buf), b + 64)
        if o.or_header_prefilter128(v1, v2):
            want.add(b)
    assert got == want and len(got) > 0


def test_fuzzed_streams_agree_with_oracle(corpus):
    """gi_test fuzz's agreement rule (gi_test.cpp:227-237): the same error class, or the same
    bytes and the same end of stream"""
    rng = refcorp.Splitmix(1)
    for name in ("f-l1", "f-l6", "f-fixed", "f-stored", "f-mixed", "f-huffman", "f-tiny"):
        stream = corpus[name][0]
        stream = stream[: 48 << 10] if len(stream) > (48 << 10) else stream
        for it in range(12):
            m = refcorp.mutate(stream, rng)
            want = oracle_inflate(m, 16 << 20)
            if want[3]:
                continue
            for seg in ((256 << 20, 0), (8 << 20, 16)):
                got = hc_inflate(m, 0, seg[0], seg[1], 16 << 20)
                assert got[0] == want[0], (name, it, seg, got[5])
                if want[0] == 0:
                    assert got[1] == want[1] and got[2] == want[2], (name, it, seg)


def test_errors():
    assert hc_inflate(b"")[0] == 2
    assert hc_inflate(bytes([0x07]))[0] == 1
    s = zlib.compress(b"xyz" * 5000, 9)[2:-4]
    assert hc_inflate(s[: len(s) - 5])[0] == 2


def test_phase1_survivors_never_adjacent():
    """the finder sizes its survivor list for every other position: its phase-1 test (BTYPE 2,
    HLIT and HDIST <= 29) never passes two adjacent bit positions, on random bits and on the
    alternating pattern that passes every other one"""
    import numpy as np
    rng = np.random.default_rng(9)

    def survivors(bits):
        n = len(bits) - 16
        p = np.arange(n)
        ok = (bits[p + 1] == 0) & (bits[p + 2] == 1)
        ok &= ~((bits[p + 4] == 1) & (bits[p + 5] == 1) & (bits[p + 6] == 1) & (bits[p + 7] == 1))
        ok &= ~((bits[p + 9] == 1) & (bits[p + 10] == 1) & (bits[p + 11] == 1) & (bits[p + 12] == 1))
        return ok
    for bits in (rng.integers(0, 2, 1 << 16), np.tile([0, 1], 1 << 15), np.tile([1, 0], 1 << 15)):
        ok = survivors(np.asarray(bits))
        assert not np.any(ok[:-1] & ok[1:])
    assert survivors(np.tile([1, 0], 1 << 15)).sum() >= (1 << 15) - 16  # the bound is reached


def test_sorted_set_membership():
    """the planner's bucketed lookup (confirming stored blocks) answers as set membership does, on
    clustered, spread, duplicated and edge values"""
    import ctypes
    import numpy as np
    h = hostcheck()
    u64p = ctypes.POINTER(ctypes.c_uint64)
    h.hc_sorted_set_has.argtypes = [u64p, ctypes.c_uint64, u64p, ctypes.c_uint64, ctypes.c_void_p]
    rng = np.random.default_rng(3)
    cases = [np.array([], np.uint64), np.array([7], np.uint64), np.array([5, 5, 5, 9], np.uint64),
             np.sort(rng.integers(0, 1 << 40, 20000, dtype=np.uint64)),
             np.sort(np.concatenate([rng.integers(1000, 1100, 500), rng.integers(1 << 33, (1 << 33) + 50, 300)]).astype(np.uint64))]
    for v in cases:
        v = np.ascontiguousarray(v)
        q = np.concatenate([v, v + 1, v - 1 if len(v) else v, rng.integers(0, 1 << 41, 5000).astype(np.uint64),
                            np.array([0, 2 ** 64 - 1], np.uint64)]).astype(np.uint64)
        out = np.zeros(len(q), np.uint8)
        h.hc_sorted_set_has(v.ctypes.data_as(u64p), len(v), q.ctypes.data_as(u64p), len(q), out.ctypes.data)
        want = np.isin(q, v).astype(np.uint8)
        assert np.array_equal(out, want)


def test_kraft_lookup_equals_direct_sum():
    """the finder's table-driven completeness test (16-bit sums, and the GPU's 8-bit saturated
    ones) equals the direct Kraft sum"""
    import ctypes
    h = hostcheck()
    h.hc_cl_complete.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    rng = random.Random(12)
    n_true = 0
    for _ in range(200000):
        v1, v2 = rng.getrandbits(64), rng.getrandbits(64)
        a, b, c = h.hc_cl_complete(v1, v2, 0), h.hc_cl_complete(v1, v2, 1), h.hc_cl_complete(v1, v2, 2)
        assert a == b == c, (hex(v1), hex(v2))
        n_true += a
    assert n_true > 100


def test_lean_header_check_equals_full_parse():
    """the finder's verify (no length array, early over-subscription exit) accepts exactly the
    positions the full parse accepts, on real streams and on random bits biased to pass the
    cheap tests"""
    import ctypes
    h = hostcheck()
    h.hc_plausible.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    o = oracle()
    n_true = 0
    for seed, level in ((1, 6), (2, 9), (3, 1)):
        s = bytes(datagen.deflate(datagen.mixed(512 << 10, seed=seed, run_max=100 << 10), level))[:60000]
        cand = np.empty(1 << 16, np.uint64)
        n = h.hc_prefilter(s, len(s), 0, len(s) * 8, cand.ctypes.data, 1 << 16)
        for b in cand[:n]:
            a = h.hc_plausible(s, len(s), int(b), 0)
            assert a == h.hc_plausible(s, len(s), int(b), 1) == h.hc_plausible(s, len(s), int(b), 2) == o.or_plausible_dynamic_header(s, len(s), int(b)), int(b)
            n_true += a
    assert n_true >= 5


@pytest.mark.parametrize("sched", [(16, 64, 8192, 32), (64, 128, 1024, 32), (8, 2048, 2048, 32), (4, 1024, 4096, 16),
                                   (4, 1024, 4096, 64), (64, 128, 1024, 64)])
def test_schedule_does_not_change_results(corpus, sched):
    """the schedule fields of gi_config_tuned (search stride, probe spacing, probe lead, lanes per
    group) move the speculative jobs and probes, never the result: every kind of stream, alone and
    as a batch, decodes to the same bytes, in_used and CRC-32 as under the default schedule"""
    h = hostcheck()
    data = datagen.mixed(3 << 20, seed=5, run_max=1 << 20)
    s6 = bytes(datagen.deflate(data, 6, 0, chunk=1 << 20))
    h.hc_set_schedule(*sched)
    try:
        for name, (stream, original) in list(corpus.items()) + [("mixed-l6", (s6, data.tobytes()))]:
            for seg in ((256 << 20, 0), (8 << 20, 64)):
                rc, out, used, *_, err = hc_inflate(stream, 0, seg[0], seg[1])
                assert rc == 0 and out == original and used == len(stream), (name, seg, err)
        streams = [s for s, _ in corpus.values()] + [s6]
        originals = [o for _, o in corpus.values()] + [data.tobytes()]
        res = hc_inflate_batch(streams, [len(o) + 64 for o in originals])
        for (r, out, used, crc), s, o in zip(res, streams, originals):
            assert r == 0 and out == o and used == len(s) and crc == zlib.crc32(o)
    finally:
        h.hc_set_schedule(0, 0, 0, 0)


@pytest.mark.parametrize("seed", [1, 3, 4, 11, 23, 42])
def test_random_schedules_and_batches(seed):
    """a random schedule (stride, spacing, lead, lanes) over random streams (sizes 20 KiB..3 MiB,
    levels 1/6/9, default/fixed/Huffman-only strategies) packed into 8 MiB batch segments as the
    library packs them (output room summed up to the segment): every stream either finishes in its
    segment with zlib's bytes and CRC-32, or is left unfinished (rc -1) and then decodes alone to
    the same bytes"""
    h = hostcheck()
    rng = random.Random(seed)
    streams = []
    for i in range(12):
        n = rng.choice([20 << 10, 200 << 10, 1 << 20, 3 << 20])
        d = datagen.mixed(n, seed=seed * 100 + i, run_min=4 << 10, run_max=128 << 10).tobytes()[:n]
        c = zlib.compressobj(rng.choice([1, 6, 9]), zlib.DEFLATED, -15, 8,
                             rng.choice([zlib.Z_DEFAULT_STRATEGY, zlib.Z_FIXED, zlib.Z_HUFFMAN_ONLY]))
        streams.append((d, c.compress(d) + c.flush()))
    sched = (rng.choice([4, 8, 16, 64]), rng.choice([64, 128, 256, 1024]), rng.choice([1024, 2048, 4096]),
             rng.choice([16, 32, 64]))
    seg_out = 8 << 20
    h.hc_set_schedule(*sched)
    try:
        left, finished = list(range(len(streams))), 0
        while left:
            take, room = [], 0
            for i in left:
                if take and room + len(streams[i][0]) + 64 > seg_out:
                    break
                take.append(i)
                room += len(streams[i][0]) + 64
            res = hc_inflate_batch([streams[i][1] for i in take], [len(streams[i][0]) + 64 for i in take],
                                   seg_out=seg_out)
            for i, (rc, out, used, crc) in zip(take, res):
                d, s = streams[i]
                if rc == -1:
                    rc, out, used, *_, err = hc_inflate(s, 0, seg_out, 64)
                    assert rc == 0 and out == d and used == len(s), (sched, i, err)
                    continue
                assert rc == 0 and out == d and used == len(s) and crc == zlib.crc32(d), (sched, i)
                finished += 1
            left = left[len(take):]
        assert finished >= len(streams) // 2, sched
    finally:
        h.hc_set_schedule(0, 0, 0, 0)


def test_threaded_backend_same_results(corpus):
    """the host build's data-parallel steps on 8 threads (the CPU baseline's speculative decoder):
    the same bytes and in_used as on one, for every kind and across tiny segments"""
    h = hostcheck()
    h.hc_set_threads.restype = None
    h.hc_set_threads.argtypes = [ctypes.c_uint]
    data = datagen.mixed(12 << 20, seed=33, run_max=2 << 20)
    s = bytes(datagen.deflate(data, 6, 0))
    try:
        h.hc_set_threads(8)
        for name, (stream, original) in list(corpus.items()) + [("mixed-12m", (s, data.tobytes()))]:
            for seg in ((256 << 20, 0), (4 << 20, 64)):
                rc, out, used, *_, err = hc_inflate(stream, 0, seg[0], seg[1])
                assert rc == 0 and out == original and used == len(stream), (name, seg, err)
    finally:
        h.hc_set_threads(1)
