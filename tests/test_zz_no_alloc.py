"""gi_rescue_headers_noalloc (the rescue's header search with every device buffer in the
library's static array, for callers that cannot allocate) against gi_rescue_headers, the entry
point whose kernels it shares: the same headers on two calls in a row, the second on a smaller
input (what the first left in the array must not show), on damaged and clean gzip files and a
raw stream; sizes past its limit are refused before any device work."""
import zlib

import numpy as np
import pytest

from gpuzip_amd import datagen
from gpuzip_amd import gpuinflate as gi
from gpuzip_amd.hipmem import DeviceBuffer
from oracle_lib import damaged

pytestmark = pytest.mark.gpu


def gz(stream, original):
    return datagen.gzip_member(stream, len(original), zlib.crc32(original))


def on_device(data):
    return DeviceBuffer.from_host(np.frombuffer(bytes(data), np.uint8))


def test_same_headers_as_the_allocating_entry_point(gpu):
    big = datagen.mixed(24 << 20, seed=31, run_max=2 << 20)
    small = datagen.text(3 << 20, seed=32)
    files = [
        ("damaged 24 MiB", damaged(gz(bytes(datagen.deflate(big, 6)), big.tobytes()), sectors=3, seed=7), 80),
        ("damaged 3 MiB, smaller", damaged(gz(bytes(datagen.deflate(small, 6)), small.tobytes()), sectors=3, seed=8), 80),
        ("clean raw 3 MiB", bytes(datagen.deflate(small, 9)), 1),
    ]
    for name, data, start in files:
        print(name, len(data), flush=True)
        assert len(data) <= 16 << 20
        want = gpu.rescue_headers(data, start)
        buf = on_device(data)
        got = gi.rescue_headers_noalloc(buf.addr, len(data), start)
        assert got == want, name
        assert len(want) > 20, name


def test_limits():
    data = bytes(17 << 20)
    buf = on_device(data)
    with pytest.raises(gi.ArgError, match="no-allocation scratch"):
        gi.rescue_headers_noalloc(buf.addr, len(data), 80)
    assert gi.rescue_headers_noalloc(buf.addr, 16, 80) == []
