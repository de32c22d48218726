"""Seeded synthetic inputs (libgzgen.so) and zlib streams of every block kind.

Content definitions live in csrc/datagen.cpp. The stream kinds follow the reference's fuzz
corpus (gpuinflate/make_fuzz_inputs.py:13-45): zlib levels 1/6/9, Z_FIXED, level 0 (stored),
mixed text/random/runs, Z_HUFFMAN_ONLY, and a tiny stream.
"""
import ctypes
import os
import random
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_gen = None

SEED_TEXT = 0xC0FFEE        # log text (config 0)
SEED_MIXED = 0xB16B00B5     # text with random runs (config 1)


def lib():
    global _gen
    if _gen is None:
        g = ctypes.CDLL(os.path.join(HERE, "lib", "libgzgen.so"))
        u64 = ctypes.c_uint64
        g.gg_text.argtypes = [u64, u64, ctypes.c_void_p, ctypes.c_int]
        g.gg_mixed.argtypes = [u64, u64, ctypes.c_double, u64, u64, ctypes.c_void_p, ctypes.c_int]
        g.gg_deflate.argtypes = [ctypes.c_void_p, u64, ctypes.c_int, ctypes.c_int, u64, ctypes.c_int, ctypes.c_void_p, u64]
        g.gg_deflate.restype = u64
        g.gg_crc32.argtypes = [ctypes.c_uint32, ctypes.c_void_p, u64, ctypes.c_int]
        g.gg_crc32.restype = ctypes.c_uint32
        _gen = g
    return _gen


def text(n, seed=SEED_TEXT, threads=0):
    """n bytes of log text as a numpy uint8 array"""
    a = np.empty(n, np.uint8)
    lib().gg_text(seed, n, a.ctypes.data, threads)
    return a


def mixed(n, seed=SEED_MIXED, random_frac=0.2, run_min=64 << 10, run_max=4 << 20, threads=0):
    """n bytes of log text interleaved with uniform random runs"""
    a = np.empty(n, np.uint8)
    lib().gg_mixed(seed, n, random_frac, run_min, run_max, a.ctypes.data, threads)
    return a


def deflate(data, level=6, strategy=0, chunk=0, threads=0):
    """raw Deflate (windowBits -15) of a numpy uint8 array or bytes; chunk > 0: parallel chunks
    primed with the previous 32 KiB and ended by sync flushes (one valid stream)"""
    a = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else data
    cap = a.size + a.size // 8 + (1 << 16) + (a.size // max(chunk, 1) + 1) * 64
    out = np.empty(cap, np.uint8)
    n = lib().gg_deflate(a.ctypes.data, a.size, level, strategy, chunk, threads, out.ctypes.data, cap)
    if n == 0 and a.size:
        raise RuntimeError("gg_deflate failed")
    return out[:n]


def crc32(data, crc=0, threads=0):
    a = np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data
    return lib().gg_crc32(crc, a.ctypes.data, a.size, threads)


def gzip_member(raw_deflate, original_len, crc, mtime=0):
    """a gzip member (RFC 1952) around a raw Deflate stream"""
    hdr = bytes([0x1F, 0x8B, 8, 0]) + int(mtime).to_bytes(4, "little") + bytes([0, 255])
    return hdr + bytes(raw_deflate) + crc.to_bytes(4, "little") + (original_len & 0xFFFFFFFF).to_bytes(4, "little")


def _zdeflate(data, level=6, strategy=zlib.Z_DEFAULT_STRATEGY):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    return c.compress(data) + c.flush()


class _Bits:
    """an LSB-first bit stream (RFC 1951 3.1.1): values go in low bit first, Huffman codes high
    bit first"""

    def __init__(self):
        self.bits = []

    def put(self, value, n):
        self.bits.extend((value >> i) & 1 for i in range(n))

    def code(self, c, n):
        self.bits.extend((c >> (n - 1 - i)) & 1 for i in range(n))

    def literal(self, b):
        """a literal in the fixed code: 8 bits for 0..143, 9 bits for 144..255"""
        if b < 144:
            self.code(0x30 + b, 8)
        else:
            self.code(0x190 + b - 144, 9)

    def align(self):
        self.bits.extend([0] * (-len(self.bits) % 8))

    def tobytes(self):
        self.align()
        return np.packbits(np.array(self.bits, np.uint8), bitorder="little").tobytes()


def _rev8(y):
    return int(f"{y:08b}"[::-1], 2)


def false_stored_stream(prefix=96 << 10, length=24000, tail=2000, seed=5):
    """(stream, data): a valid raw Deflate stream built to mislead a stored-block finder. Its first
    block is fixed-code literals (one byte each, byte-aligned after five 9-bit ones); at byte `at`
    of it, four literals read as a stored block's LEN and NLEN (LEN >= length), and LEN + 4 bytes
    further the fixed block's end-of-block code ends exactly where a real stored block (tail
    random bytes) begins, whose LEN follows one byte later. So the false candidate at `at` has a
    stored candidate right after its end, as a real stored block would. A final fixed block of 100
    literals ends the stream. Returns (stream, data, at)."""
    rng = np.random.default_rng(seed)
    ok = lambda y: 0x30 <= _rev8(y) <= 0xBF  # noqa: E731  (a byte an 8-bit fixed literal writes)
    n = length
    while not all(ok(y) for y in (n & 255, n >> 8, ~n & 255, (~n >> 8) & 255)):
        n += 1
    if n > 0xFFFF:
        raise ValueError("no LEN in range")
    w = _Bits()
    data = bytearray()

    def lit(b):
        w.literal(int(b))
        data.append(int(b))

    w.put(0, 1)  # BFINAL 0, fixed code
    w.put(1, 2)
    for b in rng.integers(144, 256, 5):
        lit(b)  # 3 + 45 bits: byte-aligned from here on
    for b in rng.integers(0, 144, prefix):
        lit(b)
    assert len(w.bits) % 8 == 0
    at = len(w.bits) // 8
    for y in (n & 255, n >> 8, ~n & 255, (~n >> 8) & 255):
        lit(_rev8(y) - 0x30)
    for b in rng.integers(0, 144, n - 2):
        lit(b)
    lit(rng.integers(144, 256))
    w.code(0, 7)  # end of block, ending at byte at + 4 + n
    assert len(w.bits) == (at + 4 + n) * 8
    w.put(0, 3)  # a stored block, its header in byte at + 4 + n
    w.align()
    w.put(tail, 16)
    w.put(~tail & 0xFFFF, 16)
    raw = rng.integers(0, 256, tail, dtype=np.uint8)
    for b in raw:
        w.put(int(b), 8)
    data += raw.tobytes()
    w.put(1, 1)  # the final block: fixed code
    w.put(1, 2)
    for b in rng.integers(0, 256, 100):
        lit(b)
    w.code(0, 7)
    return w.tobytes(), bytes(data), at
