"""Seeded damaged inputs for the rescue parity tests, modelled on the reference's own rescue
checks (tests/test_gpuzip.py:347-575, realistic.py:344-350): zeroed 512-byte sectors in a gzip
of log text with random runs (the damaged-file workload), single-block files of code and of
repetitive data with zeros over the middle, a bomb, flipped bits, CR LF transfers both ways,
several members, trailing data, a damaged header, raw streams with and without a known size and
CRC, fixed-Huffman and stored streams, a tar with a .gz inside, GNU gzip's block habit. Each case
is (name, input bytes, keyword arguments of rescue_lib.rescue)."""
import gzip
import io
import os
import random
import subprocess
import tarfile
import zlib

from gpuzip_amd import datagen
from oracle_lib import damaged

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def gz_of(data, level=6, strategy=zlib.Z_DEFAULT_STRATEGY):
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    raw = c.compress(data) + c.flush()
    return datagen.gzip_member(raw, len(data), zlib.crc32(data)), raw


def code_text():
    """60 KB of this project's own C++ sources, as the reference uses its own (test_gpuzip.py:519-523):
    a copy kept under tests/golden, so that the fixtures made from it hold while the sources change"""
    return open(os.path.join(ROOT, "tests", "golden", "code_sample.txt"), "rb").read()


def zero_middle(data, n):
    b = bytearray(data)
    mid = len(b) // 2
    b[mid:mid + n] = bytes(n)
    return bytes(b)


def flip_bits(data, k, seed):
    b = bytearray(data)
    rng = random.Random(seed)
    for _ in range(k):
        b[rng.randrange(len(b))] ^= 1 << rng.randrange(8)
    return bytes(b)


def gnu_gzip(data, level=6):
    try:
        return subprocess.run(["gzip", f"-{level}", "-n", "-c"], input=data, capture_output=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return None


def small_cases():
    """a few MiB each: the CPU parity suite (against the reference's rescue) and the GPU one"""
    cases = []
    mixed = datagen.mixed(6 << 20, seed=7, run_max=1 << 20).tobytes()
    g, raw = gz_of(mixed)
    cases.append(("mixed 6 MiB clean", g, {}))
    for seed in range(3):
        cases.append((f"mixed 6 MiB, 3 sectors zeroed (seed {seed})", damaged(g, sectors=3, seed=seed), {}))
    cases.append(("mixed, sectors zeroed, no inference", damaged(g, sectors=3, seed=1), {"no_infer": 1}))
    cases.append(("mixed, sectors zeroed, fill 255", damaged(g, sectors=3, seed=2), {"fill": 255}))
    text = datagen.text(3 << 20, seed=9).tobytes()
    gt, rawt = gz_of(text, 9)
    cases.append(("text level 9, 5 sectors zeroed", damaged(gt, sectors=5, seed=4), {}))
    cases.append(("text, truncated at two thirds", gt[: len(gt) * 2 // 3], {}))
    cases.append(("text, 0xFF run", bytes(gt[:len(gt) // 3]) + b"\xff" * 700 + bytes(gt[len(gt) // 3 + 700:]), {}))
    cases.append(("text, 8 bits flipped", flip_bits(gt, 8, "flip"), {}))
    cases.append(("text, every CR LF made LF", gt.replace(b"\r\n", b"\n"), {}))
    cases.append(("text, LF made CR LF", gt.replace(b"\n", b"\r\n"), {}))
    cases.append(("text, zero tail", gt + bytes(4096), {}))
    # one Deflate block: code, both line ends, zeros over the middle; repetitive data
    code = code_text()
    for eol, t in (("LF", code), ("CR LF", code.replace(b"\n", b"\r\n"))):
        for n in (20000, 60000):
            for zeros in (512, 64):
                cases.append((f"one block, {n // 1000} KB of code ({eol}), {zeros} zeros",
                              zero_middle(gzip.compress(t[:n], 6, mtime=0), zeros), {}))
    rep = b"gpuzip release check " * 200000
    for level in (6, 9):
        cases.append((f"repetitive 4.2 MB level {level}, 4096 zeros", zero_middle(gzip.compress(rep, level, mtime=0), 4096), {}))
    # members, trailing bytes, a damaged header
    m1, _ = gz_of(text[: 1 << 20])
    m2, _ = gz_of(mixed[: 2 << 20])
    cases.append(("two members, the second damaged", m1 + damaged(m2, sectors=2, seed=3), {}))
    cases.append(("member then garbage", m1 + b"not gzip at all" * 10, {}))
    hdr = bytearray(g)
    hdr[0:10] = bytes(10)
    cases.append(("gzip header zeroed", damaged(bytes(hdr), sectors=1, seed=8), {}))
    # raw streams, as a ZIP's entries hand them over
    rawd = damaged(raw, sectors=3, seed=5)
    cases.append(("raw, size and CRC known", rawd, {"raw": (len(mixed), zlib.crc32(mixed))}))
    cases.append(("raw, size known", rawd, {"raw": (len(mixed), None)}))
    cases.append(("raw, nothing known", rawd, {"raw": (None, None)}))
    cases.append(("raw clean, size and CRC known", raw, {"raw": (len(mixed), zlib.crc32(mixed))}))
    # the fixed code, stored blocks
    gf, _ = gz_of(text[: 2 << 20], 6, zlib.Z_FIXED)
    cases.append(("fixed Huffman, sectors zeroed", damaged(gf, sectors=3, seed=6), {}))
    gs, _ = gz_of(mixed[: 1 << 20], 0)
    cases.append(("stored, sectors zeroed", damaged(gs, sectors=3, seed=7), {}))
    # a tar holding a .gz: its block headers lie inside the outer stream's stored blocks
    inner = gzip.compress(text[: 600000], 6, mtime=0)
    buf = io.BytesIO()
    with tarfile.open(fileobj=buf, mode="w", format=tarfile.USTAR_FORMAT) as tf:
        for name, payload in (("a.txt", text[1 << 20:(1 << 20) + 300000]), ("inner.txt.gz", inner),
                              ("b.bin", mixed[: 400000])):
            ti = tarfile.TarInfo(name)
            ti.size = len(payload)
            ti.mtime = 0
            tf.addfile(ti, io.BytesIO(payload))
    tgz = gzip.compress(buf.getvalue(), 6, mtime=0)
    cases.append(("tar.gz with a .gz inside, clean", tgz, {}))
    cases.append(("tar.gz with a .gz inside, zeroed", damaged(tgz, sectors=3, seed=9), {}))
    gnu = gnu_gzip(text[: 2 << 20])
    if gnu:
        cases.append(("GNU gzip, clean", gnu, {}))
        cases.append(("GNU gzip, sectors zeroed", damaged(gnu, sectors=3, seed=10), {}))
        cases.append(("GNU gzip, 8 bits flipped", flip_bits(gnu, 8, "gnu"), {}))
    # a bomb: the limit stops it
    bomb = gzip.compress(bytes(64 << 20), 9, mtime=0)
    cases.append(("bomb, max_output 1 MiB", bomb, {"max_output": 1 << 20}))
    cases.append(("bomb, no limit", bomb, {}))
    return cases
