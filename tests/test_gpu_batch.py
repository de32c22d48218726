"""Batches (gi_batch_*): many device-resident streams through a pool of decoder contexts. Every
stream's bytes, CRC-32 and in_used match zlib, a damaged stream fails alone, and each result is
the same alone and in the mix (decoded by a single context)."""
import ctypes
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from gpuzip_amd import gpuinflate as gi
from gpuzip_amd.hipmem import DeviceBuffer
import reference_corpus as refcorp

pytestmark = pytest.mark.gpu


def raw(data, level, strategy=0):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 8, strategy)
    return c.compress(data) + c.flush()


def corpus():
    rng = np.random.default_rng(12)
    out = []
    for i in range(36):
        n = int(np.exp(rng.uniform(np.log(1 << 10), np.log(6 << 20))))
        d = datagen.mixed(n, seed=100 + i, run_min=4 << 10, run_max=256 << 10).tobytes()[:n]
        level = (1, 6, 9)[i % 3]
        strategy = zlib.Z_FIXED if i % 11 == 5 else zlib.Z_HUFFMAN_ONLY if i % 11 == 7 else 0
        out.append((d, raw(d, 0 if i % 13 == 4 else level, strategy)))
    out.append((b"", raw(b"", 6)))
    return out


def layout(streams, bad=None):
    """one input buffer (each stream 256-byte aligned) and one output buffer; Job array"""
    offs, pos = [], 0
    for _, s in streams:
        offs.append(pos)
        pos += (len(s) + 64 + 255) // 256 * 256
    host = np.zeros(pos, np.uint8)
    for (d, s), o in zip(streams, offs):
        host[o:o + len(s)] = np.frombuffer(s, np.uint8)
    if bad is not None:
        o = offs[bad]
        host[o + len(streams[bad][1]) // 3: o + len(streams[bad][1]) // 3 + 40] ^= 0x5A
    d_in = DeviceBuffer.from_host(host)
    ooffs, opos = [], 0
    for d, _ in streams:
        ooffs.append(opos)
        opos += (len(d) + 255) // 256 * 256 + 256
    d_out = DeviceBuffer(opos)
    jobs = (gi.Job * len(streams))()
    for k, ((d, s), o, oo) in enumerate(zip(streams, offs, ooffs)):
        jobs[k].d_in = d_in.addr + o
        jobs[k].in_size = len(s)
        jobs[k].d_out = d_out.addr + oo
        jobs[k].out_cap = len(d) + 256
    return d_in, d_out, jobs, ooffs


def test_batch_matches_zlib_and_single(gpu):
    streams = corpus()
    bad = 20
    d_in, d_out, jobs, ooffs = layout(streams, bad=bad)
    with gi.Batch(workers=2, segment_out_mb=64) as b:
        b.run(jobs)
        kt, nseg, nalone = b.kernel_times(reset=True)
        assert b.kernel_times() == ({}, 0, 0)
    # 37 streams in a few batch segments: a find graph, one group-decoder launch a round and a
    # resolve graph each (a stream the batch could not finish is decoded alone)
    print("batch launches", kt, nseg, nalone, flush=True)
    assert 1 <= nseg <= 4 and nalone <= 2
    assert kt["find_graph"][0] >= nseg and kt["resolve_graph"][0] >= nseg
    assert nseg <= kt["k_groups"][0] <= 8 * (nseg + nalone)
    out = d_out.to_host()
    for k, (d, s) in enumerate(streams):
        j = jobs[k]
        if k == bad:
            # damaged: an error, or (damage inside stored bytes) other bytes
            assert j.status in (gi.GI_ERR_DATA, gi.GI_ERR_TRUNCATED) or (
                j.status == gi.GI_OK and out[ooffs[k]:ooffs[k] + j.out_size].tobytes() != d), j.status
            continue
        assert j.status == gi.GI_OK, (k, j.status)
        assert j.out_size == len(d) and j.in_used == len(s) and j.crc32 == zlib.crc32(d), k
        assert out[ooffs[k]:ooffs[k] + len(d)].tobytes() == d, k
    # alone: the same streams through one context, one after another
    for k in (0, 7, 19, 33, len(streams) - 1):
        j = jobs[k]
        n, crc, st = gpu.inflate_device(j.d_in, j.in_size, j.d_out, j.out_cap)
        assert (n, crc, st.in_used) == (j.out_size, j.crc32, j.in_used), k


def test_batch_segment_each_result_as_alone(gpu):
    """one batch segment holding every kind of stream, a damaged, a truncated and an empty one:
    every job's status, bytes, CRC-32 and in_used equal the same stream decoded alone"""
    streams = corpus()[:30]
    kinds = refcorp.fuzz_corpus()
    for name in ("f-fixed", "f-stored", "f-huffman", "f-tiny"):
        s, d = kinds[name]
        streams.append((d, s))
    s, d, _ = datagen.false_stored_stream()
    streams.append((d, s))
    l6 = kinds["f-l6"][0]
    streams.append((b"", l6[: len(l6) // 2]))  # truncated
    bad = bytearray(kinds["f-l1"][0])
    bad[100] ^= 0xFF
    streams.append((b"", bytes(bad)))  # damaged
    d_in, d_out, jobs, ooffs = layout(streams)
    for k in range(len(jobs)):
        jobs[k].out_cap = max(jobs[k].out_cap, 4 << 20)
    d_out = DeviceBuffer(sum(int(jobs[k].out_cap) + 256 for k in range(len(jobs))))
    o = 0
    for k in range(len(jobs)):
        jobs[k].d_out = d_out.addr + o
        ooffs[k] = o
        o += int(jobs[k].out_cap) + 256
    with gi.Batch(workers=1) as b:
        b.run(jobs)
    out = d_out.to_host()
    got = [(jobs[k].status, jobs[k].out_size, jobs[k].crc32, jobs[k].in_used,
            out[ooffs[k]:ooffs[k] + jobs[k].out_size].tobytes()) for k in range(len(jobs))]
    for k, (d, s) in enumerate(streams):
        j = jobs[k]
        try:
            n, crc, st = gpu.inflate_device(j.d_in, j.in_size, j.d_out, j.out_cap)
            alone = (gi.GI_OK, n, crc, st.in_used, d_out.to_host(n, offset=ooffs[k]).tobytes())
        except gi.Error as e:
            alone = (e.code, 0, 0, 0, b"")
        assert got[k] == alone, (k, got[k][:4], alone[:4])
        if k < len(streams) - 2:
            assert got[k][0] == gi.GI_OK and got[k][4] == d and got[k][2] == zlib.crc32(d), k
    assert got[-2][0] == gi.GI_ERR_TRUNCATED and got[-1][0] == gi.GI_ERR_DATA


def test_serving_submit_and_wait():
    streams = corpus()[:24]
    d_in, d_out, jobs, ooffs = layout(streams)
    with gi.Batch(workers=2, segment_out_mb=64) as b:
        for k in range(len(streams)):
            b.submit(jobs[k])
        seen = set()
        while len(seen) < len(streams):
            a = b.wait(30.0)
            assert a is not None
            seen.add((a - ctypes.addressof(jobs)) // ctypes.sizeof(gi.Job))
    assert seen == set(range(len(streams)))
    out = d_out.to_host()
    for k, (d, s) in enumerate(streams):
        assert jobs[k].status == gi.GI_OK and jobs[k].crc32 == zlib.crc32(d)
        assert out[ooffs[k]:ooffs[k] + len(d)].tobytes() == d
        assert jobs[k].queued_s >= 0 and jobs[k].seconds > 0


@pytest.mark.parametrize("tuning", [
    {"chunk_kib": 16, "probe_spacing_kbit": 64, "probe_lead_bits": 8192},
    {"chunk_kib": 64, "probe_spacing_kbit": 128, "probe_lead_bits": 1024, "segment_in_max_mb": 64},
    {"lanes_per_group": 16},
    {"lanes_per_group": 32},
    {"lanes_per_group": 64, "persistent_per_cu": 10},
    {"persistent_per_cu": 3, "chunk_kib": 8, "probe_lead_bits": 2048},
])
def test_tuned_schedule_same_results(gpu, tuning):
    """a gi_config_tuned schedule (search stride, probes, group width, the persistent launch) moves
    jobs and probes, not results: every stream's bytes, size, in_used and CRC-32 as zlib's,
    batched and through one context"""
    streams = corpus()
    d_in, d_out, jobs, ooffs = layout(streams)
    with gi.Batch(workers=2, segment_out_mb=64, tuning=tuning) as b:
        assert b.run(jobs) == gi.GI_OK
    out = d_out.to_host()
    for k, (d, s) in enumerate(streams):
        j = jobs[k]
        assert j.status == gi.GI_OK and j.out_size == len(d) and j.in_used == len(s) and j.crc32 == zlib.crc32(d), k
        assert out[ooffs[k]:ooffs[k] + len(d)].tobytes() == d, k
    big = datagen.mixed(24 << 20, seed=8, run_max=2 << 20).tobytes()
    s = raw(big, 6)
    with gi.Gpu(segment_out_mb=64, tuning=tuning) as g:
        got, st = g.inflate(s)
        assert got == big and st.in_used == len(s) and g.last_crc32() == zlib.crc32(big)


def test_autotuner_on_the_device(gpu):
    """the autotuner over a small space on a small batch: its choice is one of the candidates,
    every candidate it measured decoded correctly (none rejected), and the chosen schedule
    decodes the batch to zlib's results"""
    from gpuzip_amd import autotune
    streams = corpus()[:24]
    d_in, d_out, jobs, ooffs = layout(streams)
    expected = [(len(d), len(s), zlib.crc32(d)) for d, s in streams]
    space = {"workers": [1, 2], "segment_out_mb": [64, 128], "chunk_kib": [4, 16],
             "probe_spacing_kbit": [1024, 256], "probe_lead_bits": [4096, 8192]}
    start = {"workers": 1, "segment_out_mb": 64, "chunk_kib": 4, "probe_spacing_kbit": 1024, "probe_lead_bits": 4096}
    best, score, trials = autotune.coordinate_search(autotune.measure_batch(jobs, expected, reps=1), space=space,
                                                     start=start, passes=1)
    assert score > 0 and all(s is not None for _, s in trials), trials
    assert all(best[k] in v for k, v in space.items())
    with gi.Batch(workers=best["workers"], segment_out_mb=best["segment_out_mb"], tuning=autotune.tuning_of(best)) as b:
        for j in jobs:
            j.status = -1
        assert b.run(jobs) == gi.GI_OK
    for j, (n, used, crc) in zip(jobs, expected):
        assert (j.status, j.out_size, j.in_used, j.crc32) == (gi.GI_OK, n, used, crc)
