"""The rescue's sequential decoder (gpuzip_amd/csrc/seqdec.h) against the reference decoder
(gpuinflate/inflate.cuh, built from its sources into oracle/_ref): run() field by field, output
symbols included, from every header, from arbitrary bits, in fixed and dynamic blocks, with
landing targets, small caps and stop bits, on clean, damaged and truncated streams; probe() at
many points. The rescue decides on these fields, so they must agree on damaged input too."""
import ctypes
import random
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from oracle_lib import OrResult, OrStart, damaged, ref
from rescue_lib import rc_lib

pytestmark = pytest.mark.skipif(ref() is None, reason="the reference sources are not here")

FIELDS = [k for k, _ in OrResult._fields_ if k != "resume"]


def run_both(data, start, targets, cap, first, stop):
    lib, rl = ref(), rc_lib()
    n = len(targets)
    tg = (OrStart * max(n, 1))(*targets)
    a, b = OrResult(), OrResult()
    oa = np.zeros(cap + 600, np.uint16)
    ob = np.zeros(cap + 600, np.uint16)
    lib.ref_run(data, len(data), ctypes.byref(start), tg, n, oa.ctypes.data, cap, first, stop, ctypes.byref(a))
    rl.rc_run(data, len(data), ctypes.byref(start), tg, n, ob.ctypes.data, cap, first, stop, ctypes.byref(b))
    fa = {k: getattr(a, k) for k in FIELDS}
    fb = {k: getattr(b, k) for k in FIELDS}
    ra = (a.resume.bit, a.resume.header, a.resume.mode, a.resume.final)
    rb = (b.resume.bit, b.resume.header, b.resume.mode, b.resume.final)
    if a.status in (2, 4):  # ERR_OVERFLOW, STOPPED: the resume state is defined
        assert ra == rb, (start.bit, ra, rb)
    assert fa == fb, (start.bit, start.mode, fa, fb)
    assert np.array_equal(oa[: a.out_len], ob[: b.out_len]), start.bit


def streams():
    rng = random.Random(3)
    t = datagen.text(1 << 20, seed=3).tobytes()
    m = datagen.mixed(1 << 20, seed=4, run_max=200 << 10).tobytes()
    out = []
    for data, level, strat in ((t, 6, 0), (t, 9, 0), (m, 6, 0), (t, 6, zlib.Z_FIXED), (t, 1, 0), (m, 0, 0)):
        c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strat)
        s = c.compress(data) + c.flush()
        out.append(s)
        out.append(damaged(s, sectors=3, seed=rng.randrange(1000)))
        out.append(s[: len(s) * 2 // 3])
    return out


def test_run_matches_reference_everywhere():
    from oracle_lib import oracle
    o = oracle()
    rng = random.Random(11)
    for s in streams():
        nbits = len(s) * 8
        # the dynamic headers (the rescue's targets)
        hs = np.zeros(100000, np.uint64)
        nh = o.or_find_headers(s, len(s), 1, nbits, hs.ctypes.data, len(hs))
        heads = [OrStart(int(b), int(b), 0, -1) for b in hs[: min(nh, len(hs))]]
        starts = [OrStart(0, 0, 0, -1)] + heads[:40]
        starts += [OrStart(rng.randrange(nbits), 0, 0, -1) for _ in range(40)]
        # inside blocks: a fixed-code state, and dynamic ones naming a header
        starts += [OrStart(rng.randrange(nbits), 0, 1, rng.choice([-1, 0])) for _ in range(20)]
        for h in heads[:10]:
            starts.append(OrStart(int(h.bit) + rng.randrange(1, 4000), int(h.bit), 2, -1))
        for st in starts:
            later = [h for h in heads if h.bit > st.bit][:50]
            for cap, stop, first in ((1 << 20, (1 << 64) - 1, 0), (3000, (1 << 64) - 1, 0),
                                     (1 << 20, st.bit + rng.randrange(1, 20000), 0), (1 << 20, (1 << 64) - 1, 1)):
                run_both(s, st, later, cap, first, stop)
                run_both(s, st, [], cap, first, stop)


def test_probe_matches_reference():
    lib, rl = ref(), rc_lib()
    if not hasattr(lib, "ref_probe"):
        pytest.skip("oracle/_ref built without ref_probe")
    lib.ref_probe.restype = None
    lib.ref_probe.argtypes = rl.rc_probe.argtypes
    rng = random.Random(5)
    for s in streams():
        nbits = len(s) * 8
        for _ in range(60):
            at = rng.randrange(8, nbits)
            mode = rng.choice([0, 1, 2])
            header = rng.randrange(0, at) if mode != 1 else (1 << 64) - 1
            frm = max(0, at - rng.choice([64, 4096, 20000]))
            a, b = OrStart(), OrStart()
            lib.ref_probe(s, len(s), mode, header, frm, at, ctypes.byref(a))
            rl.rc_probe(s, len(s), mode, header, frm, at, ctypes.byref(b))
            assert (a.bit, a.header, a.mode, a.final) == (b.bit, b.header, b.mode, b.final), (at, mode, header)
