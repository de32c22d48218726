"""The rescue's header search (gpuinflate/rescue.cpp:597-626, Rescuer::find_headers, with its
verify_at at rescue.cpp:549-554), as this project's decoder core runs it (the host build,
libgzhostcheck.so: plausible headers by the block finder, each verified by a decode that writes
nothing), against the reference's own code compiled from its sources (oracle/_ref/libgrref.so):
the same headers, on clean and damaged gzip files and raw streams of every kind."""
import zlib

import pytest

from gpuzip_amd import datagen
from oracle_lib import damaged, grref, hc_rescue_headers, ref_rescue_headers

needs_ref = pytest.mark.skipif(grref() is None, reason="oracle/_ref/libgrref.so not built (no reference sources)")


def gz(stream, original):
    return datagen.gzip_member(stream, len(original), zlib.crc32(original))


@needs_ref
def test_corpus_gzip_clean_and_damaged(corpus):
    for name, (stream, original) in corpus.items():
        for data in (gz(stream, original), damaged(gz(stream, original), seed=len(stream))):
            got, want = hc_rescue_headers(data, 80), ref_rescue_headers(data, 80)
            assert got == want, (name, len(got), len(want))


@needs_ref
def test_raw_streams_from_bit_one(corpus):
    for name in ("f-l1", "f-l9", "f-mixed", "f-huffman"):
        stream, _ = corpus[name]
        data = damaged(stream, sectors=2, seed=3)
        assert hc_rescue_headers(data, 1) == ref_rescue_headers(data, 1), name


@needs_ref
def test_mixed_levels_damaged():
    d = datagen.mixed(6 << 20, seed=31, run_max=1 << 20)
    for level in (1, 6, 9):
        s = bytes(datagen.deflate(d, level))
        data = damaged(gz(s, d.tobytes()), sectors=3, seed=level)
        got, want = hc_rescue_headers(data, 80), ref_rescue_headers(data, 80)
        assert got == want and len(got) > 10, level


def test_no_headers_in_noise():
    import numpy as np
    noise = np.random.default_rng(1).integers(0, 256, 1 << 16, dtype=np.uint8).tobytes()
    got = hc_rescue_headers(noise, 1)
    want = ref_rescue_headers(noise, 1)
    if want is not None:
        assert got == want
    assert len(got) <= 2
