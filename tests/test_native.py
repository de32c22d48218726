"""Host code on the CPU (no GPU): the decoding core's landing rule (a target of another kind on
a symbol boundary is passed over, never waited on); the gi_batch job queue under
ThreadSanitizer with workers, whole-batch callers and a serving loop at once; and the planner and
decoding core (the host build of csrc/deflate_core.h and csrc/plan.h) under AddressSanitizer and
UndefinedBehaviorSanitizer over every corpus kind, tiny segments and a batch segment. ROCm's clang
is used: its ThreadSanitizer runtime follows condition-variable waits that GCC 11's does not."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


def _runtime(name):
    hits = glob.glob(f"/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.{name}-x86_64.so")
    return hits[0] if hits else None


needs_clang = pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang not installed")


@needs_clang
def test_batch_queue_under_thread_sanitizer(tmp_path):
    exe = str(tmp_path / "batch_queue_test")
    subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-fsanitize=thread", "-pthread", "-o", exe,
                    os.path.join(ROOT, "tests", "native", "batch_queue_test.cpp")], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0 and "0 failures" in r.stdout, r.stdout + r.stderr[-2000:]


DRIVER = r'''
import sys, zlib
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
import numpy as np
from gpuzip_amd import datagen, workloads
from oracle_lib import hc_inflate, hc_inflate_batch
import reference_corpus as refcorp
corpus = refcorp.fuzz_corpus()
for name, (s, d) in corpus.items():
    for seg in ((256 << 20, 0), (8 << 20, 16)):
        rc, out, used, *_ , err = hc_inflate(s + b"TRAILER!", 0, seg[0], seg[1])
        assert rc == 0 and out == d and used == len(s), (name, seg, err)
streams = [s for s, _ in corpus.values()]
streams += [streams[0][:len(streams[0]) // 2], bytes([streams[1][0] ^ 0xFF]) + streams[1][1:]]
res = hc_inflate_batch(streams, [8 << 20] * len(streams))
for k, s in enumerate(streams[:len(corpus)]):
    assert res[k][0] == 0, (k, res[k][:3])
print("sanitized host build: ok")
'''


@needs_clang
@pytest.mark.skipif(_runtime("asan") is None, reason="no AddressSanitizer runtime")
def test_planner_and_core_under_address_and_ub_sanitizers(tmp_path):
    lib = str(tmp_path / "libgzhostcheck_asan.so")
    subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-shared-libsan", "-o", lib,
                    os.path.join(ROOT, "gpuzip_amd", "csrc", "hostcheck.cpp")], check=True)
    env = dict(os.environ, GZ_HOSTCHECK_LIB=lib, LD_PRELOAD=_runtime("asan"),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", DRIVER, ROOT], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "ok" in r.stdout, (r.stdout + r.stderr)[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_landing_rule_passes_over_targets_of_another_kind(tmp_path):
    """a decode whose landing list has a block start, or another header's start, exactly on one of
    its symbol boundaries runs on to the stream's end with the same output (each case in a child
    process with an alarm: a decode that never returns fails its case)"""
    exe = str(tmp_path / "landing_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "gpuzip_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "native", "landing_test.cpp"), "-lz"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "0 failures" in r.stdout, r.stdout[-4000:] + r.stderr[-2000:]
