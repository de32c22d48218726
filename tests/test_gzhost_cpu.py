"""The gzip member driver (gpuzip_amd/csrc/gzhost.cpp) on the CPU: built with a zlib stand-in for
gi_inflate (tests/native/fake_gpuinflate.cpp), so its member walk, trailer checks and failure
classes (src/gz.rs:154-264: Corrupt, Retry, Stopped) are tested without a GPU, with faults acted
out on demand: a device error and a delivered buffer changed after the GPU's CRC was taken are
both GZ_RETRY (the caller decodes on the CPU), a wrong GPU CRC too. Runs of BGZF members go
through gi_inflate_many, as the reference's 301-member bgzip test file (tests/test_gpuzip.py:
1008-1040) does here."""
import ctypes
import gzip
import os
import subprocess
import zlib

import pytest

from gpuzip_amd import gpuinflate as gi
from gpuzip_amd import gz

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def fake(tmp_path_factory):
    lib = str(tmp_path_factory.mktemp("gzfake") / "libgzhost_fake.so")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"), "-o", lib,
                    os.path.join(ROOT, "gpuzip_amd", "csrc", "gzhost.cpp"),
                    os.path.join(ROOT, "tests", "native", "fake_gpuinflate.cpp"), "-lz"], check=True)
    return lib


def decode(fake_path, data, fault=None, stop_after=None):
    """gz_decode in a child process (FAKE_GI_FAULT is read per call): (rc, output, report fields)"""
    script = r'''
import ctypes, sys, json
sys.path.insert(0, sys.argv[1])
from gpuzip_amd import gpuinflate as gi
from gpuzip_amd import gz
lib = ctypes.CDLL(sys.argv[2])
lib.fake_gi_create.restype = ctypes.c_void_p
lib.gz_decode.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, gi.SINK, ctypes.c_void_p, ctypes.POINTER(gz.Report)]
data = sys.stdin.buffer.read()
out = bytearray()
stop = int(sys.argv[3])
def sink(u, p, n):
    out.extend(ctypes.string_at(p, n))
    return 1 if stop and len(out) >= stop else 0
cb = gi.SINK(sink)
rep = gz.Report()
ctx = lib.fake_gi_create()
rc = lib.gz_decode(ctx, data, len(data), cb, None, ctypes.byref(rep))
lib.fake_gi_calls.argtypes = [ctypes.c_void_p, ctypes.c_int]
sys.stdout.buffer.write(json.dumps({"rc": rc, "members": rep.members, "out": rep.out, "error": rep.error.decode(),
                                    "info": rep.info.decode(), "note": rep.note.decode(), "len": len(out),
                                    "crc": __import__("zlib").crc32(bytes(out)), "alone_calls": lib.fake_gi_calls(ctx, 0),
                                    "many_calls": lib.fake_gi_calls(ctx, 1)}).encode())
'''
    env = dict(os.environ)
    env.pop("FAKE_GI_FAULT", None)
    if fault:
        env["FAKE_GI_FAULT"] = fault
    r = subprocess.run(["python3", "-c", script, ROOT, fake_path, str(stop_after or 0)], input=bytes(data),
                       capture_output=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    import json
    return json.loads(r.stdout)


def members(*payloads, **kw):
    return b"".join(gzip.compress(p, mtime=0, **kw) for p in payloads)


PAYLOADS = [bytes(range(256)) * 300, b"hello gzip " * 5000, os.urandom(100000)]


def test_members_decode_and_check(fake):
    data = members(*PAYLOADS)
    r = decode(fake, data)
    want = b"".join(PAYLOADS)
    assert r["rc"] == gz.GZ_OK and r["members"] == 3 and r["out"] == len(want) == r["len"], r
    assert r["crc"] == zlib.crc32(want)


def test_trailing_zeros_and_garbage(fake):
    r = decode(fake, members(PAYLOADS[0]) + bytes(100))
    assert r["rc"] == gz.GZ_OK and r["info"] == "100 zero bytes after the last member", r
    r = decode(fake, members(PAYLOADS[0]) + b"junk")
    assert r["rc"] == gz.GZ_OK and r["note"] == "4 bytes after the last member are not gzip data; ignored", r


def test_device_fault_is_retry(fake):
    r = decode(fake, members(*PAYLOADS), fault="gpu:2")
    assert r["rc"] == gz.GZ_RETRY and r["error"] == "GPU error: injected device fault", r


def test_changed_delivery_is_retry(fake):
    """the GPU's CRC right, the bytes the sink received not: caught by the host CRC"""
    r = decode(fake, members(*PAYLOADS), fault="flip:3")
    assert r["rc"] == gz.GZ_RETRY and r["error"] == "GPU output failed the CRC check", r
    assert r["members"] == 2


def test_wrong_gpu_crc_is_retry(fake):
    r = decode(fake, members(*PAYLOADS), fault="crc:1")
    assert r["rc"] == gz.GZ_RETRY and r["error"] == "GPU output failed the CRC check", r


def test_damage_and_truncation_are_corrupt(fake):
    good = members(PAYLOADS[1])
    r = decode(fake, good[:len(good) // 2])
    assert r["rc"] == gz.GZ_CORRUPT, r
    r = decode(fake, good[:-3])
    assert r["rc"] == gz.GZ_CORRUPT and r["error"] == "the file ends before a member's trailer", r
    bad_size = bytearray(good)
    bad_size[-1] ^= 1
    r = decode(fake, bytes(bad_size))
    assert r["rc"] == gz.GZ_RETRY and r["error"] == "GPU output failed the size check", r
    r = decode(fake, b"")
    assert r["rc"] == gz.GZ_CORRUPT and r["error"] == "no gzip member", r
    r = decode(fake, b"\x1f\x8b\x09" + bytes(20))
    assert r["rc"] == gz.GZ_CORRUPT and r["error"] == "compression method 9 is not Deflate", r


def test_stopped_by_sink(fake):
    r = decode(fake, members(*PAYLOADS), stop_after=1000)
    assert r["rc"] == gz.GZ_STOPPED and r["error"] == "stopped by the caller", r


def bgzf_member(data):
    """one BGZF member as bgzip writes it: the BC extra subfield holds the member's size less one"""
    import struct
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    z = c.compress(data) + c.flush()
    head = struct.pack("<BBBBIBBH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6) + b"BC" + struct.pack("<HH", 2, 18 + len(z) + 8 - 1)
    return head + z + struct.pack("<II", zlib.crc32(data), len(data))


def bgzf_blocks(n=300, seed=11):
    import random
    rnd = random.Random(seed)
    words = ["alpha", "beta", "gamma", "delta", "read", "quality", "chr1", "chr2", "ACGT", "TTGA"]
    return [" ".join(rnd.choice(words) for _ in range(9000)).encode()[:65280] for _ in range(n)]


def test_bgzf_members_decode_together(fake):
    """301 members (300 of 64 KB and bgzip's empty one at the end): one gi_inflate_many call, no
    member decoded by itself, every trailer checked"""
    blocks = bgzf_blocks()
    data = b"".join(bgzf_member(b) for b in blocks) + bgzf_member(b"")
    r = decode(fake, data)
    want = b"".join(blocks)
    assert r["rc"] == gz.GZ_OK and r["members"] == 301 and r["len"] == len(want) and r["crc"] == zlib.crc32(want)
    assert r["many_calls"] == 1 and r["alone_calls"] == 0


def test_bgzf_damaged_member_is_reported_as_alone(fake):
    """a damaged member in a run: the members before it reach the sink, it is decoded by itself
    and reported as the one-member path reports it"""
    blocks = bgzf_blocks(20)
    ms = [bgzf_member(b) for b in blocks]
    bad = bytearray(ms[7])
    bad[30:60] = bytes(30)
    data = b"".join(ms[:7]) + bytes(bad) + b"".join(ms[8:])
    r = decode(fake, data)
    alone = decode(fake, bytes(bad))
    assert r["rc"] == gz.GZ_CORRUPT and r["members"] == 7 and r["error"] == alone["error"] and alone["rc"] == gz.GZ_CORRUPT
    assert r["len"] == sum(len(b) for b in blocks[:7]) and r["alone_calls"] == 1


def test_bgzf_wrong_gpu_crc_is_decoded_again(fake):
    """a batched member whose GPU CRC fails its trailer goes through gi_inflate by itself; a new
    run starts after it (the stand-in spoils the sixth stream of every batched call: members 5
    and 11 of the 12 are decoded alone)"""
    blocks = bgzf_blocks(12)
    data = b"".join(bgzf_member(b) for b in blocks)
    r = decode(fake, data, fault="mcrc:5")
    want = b"".join(blocks)
    assert r["rc"] == gz.GZ_OK and r["members"] == 12 and r["crc"] == zlib.crc32(want)
    assert r["alone_calls"] == 2 and r["many_calls"] == 2


def test_bgzf_device_fault_is_retry(fake):
    data = b"".join(bgzf_member(b) for b in bgzf_blocks(4))
    r = decode(fake, data, fault="many:1")
    assert r["rc"] == gz.GZ_RETRY and "GPU error" in r["error"]


def test_bgzf_runs_around_plain_members(fake):
    """BGZF runs, a plain member between them, a lone BGZF member, trailing zeros"""
    blocks = bgzf_blocks(9)
    plain = gzip.compress(PAYLOADS[1], mtime=0)
    data = (b"".join(bgzf_member(b) for b in blocks[:4]) + plain + b"".join(bgzf_member(b) for b in blocks[4:8]) +
            plain + bgzf_member(blocks[8]) + bytes(64))
    r = decode(fake, data)
    want = b"".join(blocks[:4]) + PAYLOADS[1] + b"".join(blocks[4:8]) + PAYLOADS[1] + blocks[8]
    assert r["rc"] == gz.GZ_OK and r["members"] == 11 and r["crc"] == zlib.crc32(want) and r["len"] == len(want)
    assert r["many_calls"] == 2 and r["alone_calls"] == 3 and "zero bytes" in r["info"]
