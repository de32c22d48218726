"""Synthetic request streams for the batched and served measurements: raw Deflate streams whose
content is the log-text / random-run mix of the single-stream benchmark, at zlib levels 1, 6 and
9 in equal shares.

  batch_sizes(n, seed)      n sizes log-uniform on [1 KiB, 64 MiB] (the batched workload)
  serve_sizes(n, seed)      n sizes lognormal, median 2 MiB, sigma 1.5, capped at 512 MiB
  make_streams(sizes, ...)  the streams: slices of one seeded content pool, each compressed on
                            the host's threads; with their CRC-32s
"""
import os
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import datagen

SEED_SIZES = 0xE3
SEED_LEVELS = 0xF3
SEED_POOL = 0xB1
KIB, MIB = 1 << 10, 1 << 20


def batch_sizes(n, seed=SEED_SIZES):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(KIB), np.log(64 * MIB), n)).astype(np.int64)


def serve_sizes(n, seed=SEED_SIZES):
    rng = np.random.default_rng(seed)
    s = rng.lognormal(np.log(2 * MIB), 1.5, n)
    return np.clip(s, KIB, 512 * MIB).astype(np.int64)


def levels(n, seed=SEED_LEVELS):
    rng = np.random.default_rng(seed)
    lv = np.array([1, 6, 9] * ((n + 2) // 3))[:n]
    rng.shuffle(lv)
    return lv


def content_pool(nbytes, seed=SEED_POOL):
    """the mix the streams are cut from: log text with 20% random runs of 64 KiB-4 MiB"""
    return datagen.mixed(nbytes, seed=seed)


def make_streams(sizes, pool, threads=None, seed=SEED_POOL):
    """[(raw Deflate bytes, output size, CRC-32)] for each size: a slice of pool at a seeded offset
    (a slice longer than the pool wraps around), compressed at its level"""
    rng = np.random.default_rng(seed + 1)
    lv = levels(len(sizes))
    starts = rng.integers(0, pool.size, len(sizes))
    mv = memoryview(pool)

    def one(k):
        n, a = int(sizes[k]), int(starts[k])
        if a + n <= pool.size:
            d = mv[a:a + n]
        else:
            d = bytes(mv[a:]) + bytes(np.resize(pool, n - (pool.size - a)))
        c = zlib.compressobj(int(lv[k]), zlib.DEFLATED, -15)
        s = c.compress(d) + c.flush()
        return s, n, zlib.crc32(d)

    with ThreadPoolExecutor(threads or min(16, os.cpu_count() or 8)) as ex:
        return list(ex.map(one, range(len(sizes))))
