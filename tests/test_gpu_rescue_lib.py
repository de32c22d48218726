"""The GPU rescue (libgzrescue.so: the rescue's decisions over GPU header search, decode and
provenance) against the reference's own rescue (gpuinflate/rescue.cpp, built from its sources into
oracle/_ref/libgrref.so) and the golden fixtures made from it: return code, output bytes, mask,
statistics, pieces and notes all equal, on the seeded damaged inputs of tests/rescue_cases.py,
on a 64 MiB damaged file of log text and random runs, and on a decode that runs past the 32 Mi
symbols after which the reference resumes a decode."""
import json
import os
import time

import numpy as np
import pytest

import rescue_cases
import rescue_lib
from gpuzip_amd import datagen
from oracle_lib import damaged

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "rescue_golden.json")))
CASES = {name: (data, kw) for name, data, kw in rescue_cases.small_cases()}


def product():
    lib = rescue_lib.product()
    if lib is None:
        pytest.fail("gpuzip_amd/lib/libgzrescue.so is not built (python -m gpuzip_amd.build)")
    return lib


def against_reference(data, **kw):
    ph = {}
    t = time.time()
    got = rescue_lib.rescue(product(), data, phases=ph, **kw)
    print(f"  GPU rescue {time.time() - t:.3f} s; " + ", ".join(f"{k} {v:.3f}" for k, v in ph.items()), flush=True)
    ref = rescue_lib.reference()
    if ref is not None:
        t = time.time()
        want = rescue_lib.rescue(ref, data, **kw)
        print(f"  reference rescue {time.time() - t:.3f} s", flush=True)
        rescue_lib.assert_same(got, want)
    return got


@pytest.mark.parametrize("name", sorted(CASES))
def test_case_matches_reference(name):
    print("case", name, flush=True)
    data, kw = CASES[name]
    got = against_reference(data, **kw)
    assert rescue_lib.summary(got) == GOLDEN[name]


def test_damaged_64mib_text_and_random_runs():
    """three zeroed 512-byte sectors in a 64 MiB file of log text with random runs, one zlib call
    at level 6 (the damaged-file workload at a sixteenth of its size)"""
    print("case 64 MiB mixed, 3 sectors zeroed", flush=True)
    d = datagen.mixed(64 << 20)
    g = datagen.gzip_member(datagen.deflate(d, 6), d.size, datagen.crc32(d))
    got = against_reference(damaged(g, 3, seed=2))
    assert got[0] == 0 and len(got[1]) == d.size and got[3]["damaged"] >= 1


def test_decode_past_the_resume_point():
    """a decode from the last header before 40 MiB of stored blocks runs past 32 Mi symbols: the
    reference resumes it there (counting a stored block that did not fit twice), and so must this"""
    print("case 42 MiB with a 40 MiB stored stretch", flush=True)
    text = datagen.text(1 << 20, seed=3)
    rnd = np.random.default_rng(5).integers(0, 256, 40 << 20, dtype=np.uint8)
    d = np.concatenate([text, rnd, datagen.text(1 << 20, seed=4)])
    g = datagen.gzip_member(datagen.deflate(d, 6), d.size, datagen.crc32(d))
    for seed in (1, 2):
        got = against_reference(damaged(g, 2, seed=seed))
        assert got[0] == 0
    got = against_reference(g)
    assert got[1] == d.tobytes() and got[3]["crc_check"] == 1


def test_host_binding():
    """gpuzip_amd.rescue (the reference loader's interface, src/rescue.rs:179-262): the same
    result through its sink; a sink that returns False stops the rescue with an error"""
    from gpuzip_amd import rescue as gr
    print("case host binding", flush=True)
    data, _ = CASES["mixed 6 MiB, 3 sectors zeroed (seed 0)"]
    out, mask = bytearray(), bytearray()

    def sink(d, m):
        out.extend(d)
        mask.extend(m)
        return True

    rep = gr.get().rescue(data, 0, sink, max_output=0)
    want = rescue_lib.rescue(product(), data)
    assert bytes(out) == want[1] and bytes(mask) == want[2]
    assert rep.stats == want[3] and rep.notes == want[5] and len(rep.pieces) == len(want[4])
    assert "k_rescue_decode" in rep.kernels and "find headers" in rep.phases
    with pytest.raises(gr.RescueError, match="stopped"):
        gr.get().rescue(data, 0, lambda d, m: False)
    raw = rescue_cases.gz_of(b"x" * 1000)[1]
    rep = gr.get().rescue_raw(raw, (1000, None), 0, sink=None)
    assert rep.stats["output_size"] == 1000
