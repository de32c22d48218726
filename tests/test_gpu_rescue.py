"""The rescue's header search on the GPU (gi_rescue_headers: the block finder, then a decode
that writes nothing from every plausible header, rescue.cpp:549-626) against the same search by
the host build of the decoder core and, where oracle/_ref was built, by the reference's own
rescue.cpp: the same headers on clean and damaged gzip files and raw streams."""
import zlib

import pytest

from gpuzip_amd import datagen
from oracle_lib import damaged, hc_rescue_headers, ref_rescue_headers

pytestmark = pytest.mark.gpu


def gz(stream, original):
    return datagen.gzip_member(stream, len(original), zlib.crc32(original))


def check(gpu, data, start):
    got = gpu.rescue_headers(data, start)
    assert got == hc_rescue_headers(data, start)
    want = ref_rescue_headers(data, start)
    if want is not None:
        assert got == want
    return got


def test_corpus_clean_and_damaged(gpu, corpus):
    for name, (stream, original) in corpus.items():
        print("kind", name, flush=True)
        check(gpu, gz(stream, original), 80)
        check(gpu, damaged(gz(stream, original), seed=len(stream)), 80)
        check(gpu, damaged(stream, sectors=2, seed=3), 1)


def test_damaged_32mib_level6(gpu):
    d = datagen.mixed(32 << 20, seed=77, run_max=2 << 20)
    s = bytes(datagen.deflate(d, 6))
    hs = check(gpu, damaged(gz(s, d.tobytes()), sectors=3, seed=5), 80)
    assert len(hs) > 100
