"""The rescue's decisions (gpuzip_amd/csrc/rescue_core.h) built for the CPU over a CPU backend
(librescuecheck.so), against the reference's own rescue (gpuinflate/rescue.cpp via
oracle/_ref/libgrref.so, or its results kept in tests/golden/rescue_golden.json): code, output,
mask, statistics, pieces and notes, all exactly equal, on the damaged inputs of
tests/rescue_cases.py. The same cases run through the GPU build in test_gpu_rescue.py."""
import json
import os
import zlib

import numpy as np
import pytest

import rescue_cases
import rescue_lib
from gpuzip_amd import datagen
from oracle_lib import damaged

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "rescue_golden.json")))
CASES = {name: (data, kw) for name, data, kw in rescue_cases.small_cases()}


def hostbuild():
    lib = rescue_lib.hostbuild()
    if lib is None:
        pytest.fail("gpuzip_amd/lib/librescuecheck.so is not built (python -m gpuzip_amd.build)")
    return lib


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_build_matches_reference(name):
    data, kw = CASES[name]
    got = rescue_lib.rescue(hostbuild(), data, **kw)
    ref = rescue_lib.reference()
    if ref is not None:
        rescue_lib.assert_same(got, rescue_lib.rescue(ref, data, **kw))
    assert rescue_lib.summary(got) == GOLDEN[name]


def test_golden_covers_every_case():
    assert sorted(GOLDEN) == sorted(CASES) or set(CASES) < set(GOLDEN)


@pytest.mark.skipif(rescue_lib.reference() is None, reason="reference rescue not built here")
def test_golden_is_the_reference():
    """the committed fixtures still describe the reference library built from its sources"""
    for name in ("mixed 6 MiB, 3 sectors zeroed (seed 0)", "one block, 20 KB of code (CR LF), 512 zeros", "raw, size known"):
        data, kw = CASES[name]
        assert rescue_lib.summary(rescue_lib.rescue(rescue_lib.reference(), data, **kw)) == GOLDEN[name]


def test_exact_bytes_are_true():
    """what the rescue calls exact was decoded from bits that decode cleanly: the output before
    the first byte that is not exact is the original's start; when exact bytes elsewhere differ
    from the original (a zeroed sector can decode in step, as a run of valid codes), the rescue
    has said the CRC could not vouch for them; unknown bytes hold the fill; the mask's classes
    agree with the statistics; on the undamaged file everything is exact and the CRC holds"""
    orig = datagen.mixed(4 << 20, seed=21, run_max=1 << 20).tobytes()
    c = zlib.compressobj(6, zlib.DEFLATED, -15)
    raw = c.compress(orig) + c.flush()
    g = datagen.gzip_member(raw, len(orig), zlib.crc32(orig))
    rc, out, mask, stats, _, _ = rescue_lib.rescue(hostbuild(), g)
    assert rc == 0 and out == orig and set(mask) == {EXACT}
    assert stats["crc_check"] == 1 and stats["exact_bytes"] == len(orig)
    for seed in range(4):
        rc, out, mask, stats, _, notes = rescue_lib.rescue(hostbuild(), damaged(g, sectors=3, seed=seed), fill=0x5A)
        assert rc == 0 and len(out) == len(orig)
        o, m, x = np.frombuffer(out, np.uint8), np.frombuffer(mask, np.uint8), np.frombuffer(orig, np.uint8)
        inexact = np.nonzero(m != EXACT)[0]
        assert len(inexact) > 0 and stats["damaged"] > 0
        assert out[:inexact[0]] == orig[:inexact[0]]
        if np.any((o != x) & (m == EXACT)):
            assert stats["crc_check"] != 1 and any("not vouched for" in n or "would not show" in n for n in notes)
        assert np.all(o[m == UNKNOWN] == 0x5A)
        counts = np.bincount(m, minlength=4)
        assert len(counts) == 4
        assert counts[EXACT] == stats["exact_bytes"] and counts[UNKNOWN] == stats["unknown_bytes"]
        assert counts[SUSPECT] == stats["suspect_bytes"]
        # the statistic also counts inferred bytes that lie in suspect blocks, which keep that mask
        assert 0 < counts[INFERRED] <= stats["inferred_bytes"]


UNKNOWN, EXACT, SUSPECT, INFERRED = 0, 1, 2, 3


def test_crc_combine():
    """the host's CRC-32 combine (host_crc.h, after zlib's crc32_combine), which joins the GPU's
    per-chunk CRCs of a member: crc(A B) from crc(A), crc(B) and |B|, against zlib on random splits"""
    import ctypes
    import random
    lib = ctypes.CDLL(os.path.join(os.path.dirname(HERE), "gpuzip_amd", "lib", "librescuecheck.so"))
    lib.rc_crc32_combine.restype = ctypes.c_uint32
    lib.rc_crc32_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
    rng = random.Random(7)
    for _ in range(200):
        a = rng.randbytes(rng.randrange(0, 3000))
        b = rng.randbytes(rng.choice([0, 1, 1023, 1024, 65536, rng.randrange(0, 200000)]))
        assert lib.rc_crc32_combine(zlib.crc32(a), zlib.crc32(b), len(b)) == zlib.crc32(a + b)
