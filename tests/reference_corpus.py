"""The reference's fuzz corpus and damage model, for tests: the eight stream kinds of
make_fuzz_inputs.py:18-45 built from this package's log text, and gi_test's splitmix64 generator
and mutations (gi_test.cpp:99-108, 197-218), kept alike so the damaged streams are the kinds the
reference's own fuzzing makes. Test infrastructure, not part of the product package."""
import random
import zlib

from gpuzip_amd.datagen import text


def _zdeflate(data, level=6, strategy=zlib.Z_DEFAULT_STRATEGY):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    return c.compress(data) + c.flush()


def fuzz_corpus(text_bytes=None):
    """The eight stream kinds of the reference's fuzz corpus (make_fuzz_inputs.py:18-45), built
    the same way from this package's log text; returns {name: (deflate bytes, original)}."""
    t = bytes(text(2 << 20)) if text_bytes is None else text_bytes
    rng = random.Random(7)
    mix = bytearray()
    while len(mix) < (1 << 20):
        k = rng.randrange(3)
        n = rng.randrange(1, 40000)
        if k == 0:
            mix += rng.randbytes(n)
        elif k == 1:
            a = rng.randrange(len(t) - n)
            mix += t[a:a + n]
        else:
            mix += bytes([rng.randrange(256)]) * n
    mix = bytes(mix)
    tiny = b"hello, hello, hello world\n" * 3
    src = {
        "f-l1": (t[: 1 << 20], 1, zlib.Z_DEFAULT_STRATEGY),
        "f-l6": (t, 6, zlib.Z_DEFAULT_STRATEGY),
        "f-l9": (t[: 1 << 20], 9, zlib.Z_DEFAULT_STRATEGY),
        "f-fixed": (t[: 1 << 20], 6, zlib.Z_FIXED),
        "f-stored": (t[: 512 << 10], 0, zlib.Z_DEFAULT_STRATEGY),
        "f-mixed": (mix, 6, zlib.Z_DEFAULT_STRATEGY),
        "f-huffman": (t[: 1 << 20], 6, zlib.Z_HUFFMAN_ONLY),
        "f-tiny": (tiny, 6, zlib.Z_DEFAULT_STRATEGY),
    }
    return {k: (_zdeflate(d, lv, st), d) for k, (d, lv, st) in src.items()}


class Splitmix:
    """gi_test.cpp:99-108's generator, for damage that reproduces across machines"""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def below(self, n):
        return self.next() % n if n else 0


def mutate(stream, rng):
    """One damaged copy of a stream, with gi_test's mutations (gi_test.cpp:197-218): bit flips,
    overwritten bytes and spans, truncation, copied spans."""
    m = bytearray(stream)
    nmut = 1 + rng.below(4)
    for _ in range(nmut):
        if not m:
            break
        k = rng.below(6)
        if k in (0, 1):
            m[rng.below(len(m))] ^= 1 << rng.below(8)
        elif k == 2:
            m[rng.below(len(m))] = rng.next() & 0xFF
        elif k == 3:
            a = rng.below(len(m))
            n = min(1 + rng.below(64), len(m) - a)
            for i in range(n):
                m[a + i] = rng.next() & 0xFF
        elif k == 4:
            del m[rng.below(len(m)) + 1:]
        else:
            a, b = rng.below(len(m)), rng.below(len(m))
            n = min(1 + rng.below(4096), len(m) - a, len(m) - b)
            m[b:b + n] = bytes(m[a:a + n])
    return bytes(m)
