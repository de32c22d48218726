"""Serving in the library (gi_serve_*, the continuous batching the batched form calls for):
gzip-member requests from host memory admitted by an HBM budget, decoded by the batch workers
folded into their next batch segments. A request decodes to the same bytes alone and when it
arrives while dozens of others are on the device; admitted requests never hold more device
memory than the budget; requests too large, damaged or with a wrong trailer get their codes."""
import gzip
import random
import threading
import time
import zlib

import pytest

from gpuzip_amd import datagen
from gpuzip_amd import gpuinflate as gi

pytestmark = pytest.mark.gpu


def member(n, seed, level=6):
    d = datagen.mixed(n, seed=seed, run_max=1 << 20).tobytes() if n > 1024 else bytes(n)
    return gzip.compress(d, level, mtime=0), d


def test_alone_and_in_the_mix():
    rng = random.Random(3)
    payloads = [member(rng.choice([1 << 10, 64 << 10, 1 << 20, 4 << 20, 12 << 20]), seed=s, level=rng.choice([1, 6, 9]))
                for s in range(48)]
    x, xd = member(8 << 20, seed=1000)
    budget = 96 << 20
    with gi.Server(budget, workers=2) as srv:
        print("alone", flush=True)
        srv.submit(x, len(xd), tag="x")
        r, alone, _ = srv.wait(60)
        assert r.status == gi.GI_OK and alone == xd and r.crc32 == zlib.crc32(xd)
        # the mix: arrivals in bursts of a dozen from another thread; x again in the middle of one
        results, submitted, lock = {}, [0], threading.Lock()
        x_in_flight = []

        ready = [gi.Server.prepare(g, len(d), tag=i) for i, (g, d) in enumerate(payloads)]
        x_again = gi.Server.prepare(x, len(xd), tag="x")

        def arrive():
            for i in range(len(payloads)):
                srv.submit(ready[i])
                with lock:
                    submitted[0] += 1
                    pending = submitted[0] - len(results)
                if i == 12 + 6:  # in the middle of the second burst
                    x_in_flight.append(pending)
                    srv.submit(x_again)
                if i % 12 == 11:  # bursts of a dozen arrivals
                    time.sleep(rng.uniform(0.005, 0.02))

        print("mix", flush=True)
        t = threading.Thread(target=arrive)
        t.start()
        deadline = time.time() + 120
        while len(results) < len(payloads) + 1 and time.time() < deadline:
            got = srv.wait(1.0)
            if got is None:
                continue
            r, out, tag = got
            with lock:
                results[tag] = (r.status, out, r.crc32, r.admitted_s, r.done_s)
        t.join()
        now, peak = srv.in_use()
    assert len(results) == len(payloads) + 1
    assert x_in_flight and x_in_flight[0] >= 4, x_in_flight
    assert results["x"][0] == gi.GI_OK and results["x"][1] == alone
    for i, (g, d) in enumerate(payloads):
        st, out, crc, adm, done = results[i]
        assert st == gi.GI_OK and out == d and crc == zlib.crc32(d), i
    assert now == 0 and 0 < peak <= budget
    # 48 requests of up to 12 MiB against a 96 MiB budget: some waited to be admitted
    assert max(v[3] for v in results.values()) > 0
    print(f"peak device bytes {peak}, slowest admission {max(v[3] for v in results.values()) * 1e3:.1f} ms", flush=True)


def test_rejections_and_damage():
    g, d = member(2 << 20, seed=5)
    bad = bytearray(g)
    bad[len(bad) // 2] ^= 0x55
    wrong_crc = bytearray(g)
    wrong_crc[-8] ^= 1
    big, bigd = member(40 << 20, seed=6)
    with gi.Server(32 << 20, workers=1) as srv:
        srv.submit(big, len(bigd), tag="big")
        srv.submit(b"not a gzip member at all", 100, tag="header")
        srv.submit(bytes(bad), len(d), tag="damaged")
        srv.submit(bytes(wrong_crc), len(d), tag="trailer")
        srv.submit(g, len(d), tag="good")
        got = {}
        for _ in range(5):
            r, out, tag = srv.wait(60)
            got[tag] = (r.status, out, r.error.decode())
    assert got["big"][0] == gi.GI_ERR_ARG and "budget" in got["big"][2]
    assert got["header"][0] == gi.GI_ERR_DATA and "not a gzip member" in got["header"][2]
    assert got["damaged"][0] in (gi.GI_ERR_DATA, gi.GI_ERR_TRUNCATED, gi.GI_ERR_ARG)
    assert got["trailer"][0] == gi.GI_ERR_DATA and "trailer" in got["trailer"][2]
    assert got["good"][0] == gi.GI_OK and got["good"][1] == d


def test_poisson_arrivals_each_as_alone():
    """requests arriving on a seeded Poisson trace (a thread submits each at its arrival time)
    while others decode: every request's status, bytes, CRC-32 and output size equal those of the
    same request sent alone to an idle server, a damaged member and an empty one included"""
    rng = random.Random(11)
    reqs = [member(int(2 ** rng.uniform(10, 23.5)), seed=500 + s, level=rng.choice([1, 6, 9])) for s in range(64)]
    g, d = member(3 << 20, seed=77)
    bad = bytearray(g)
    bad[len(bad) // 3:len(bad) // 3 + 24] = b"\x00" * 24
    reqs += [(bytes(bad), d), member(0, seed=0)]
    budget = 128 << 20
    with gi.Server(budget, workers=2) as srv:
        alone = []
        for k, (m, dd) in enumerate(reqs):
            srv.submit(m, len(dd) + 64, tag=k)
            r, out, _ = srv.wait(60)
            alone.append((r.status, out, r.crc32, r.out_size))
        assert alone[-2][0] != gi.GI_OK and alone[-1][:2] == (gi.GI_OK, b"")
        assert all(a[0] == gi.GI_OK and a[1] == dd for a, (_, dd) in zip(alone[:-2], reqs[:-2]))
        # the mix: 66 arrivals at 400 a second on average, in a shuffled order
        order = list(range(len(reqs)))
        rng.shuffle(order)
        ready = [gi.Server.prepare(reqs[k][0], len(reqs[k][1]) + 64, tag=k) for k in order]
        gaps = [rng.expovariate(400.0) for _ in order]
        results, peak_pending, lock = {}, [0], threading.Lock()

        def arrive():
            t = time.perf_counter()
            for i, (p, gap) in enumerate(zip(ready, gaps)):
                t += gap
                time.sleep(max(0.0, t - time.perf_counter()))
                srv.submit(p)
                with lock:
                    peak_pending[0] = max(peak_pending[0], i + 1 - len(results))

        th = threading.Thread(target=arrive)
        th.start()
        deadline = time.time() + 120
        while len(results) < len(reqs) and time.time() < deadline:
            got = srv.wait(1.0)
            if got is None:
                continue
            r, out, k = got
            with lock:
                results[k] = (r.status, out, r.crc32, r.out_size)
        th.join()
        now, peak = srv.in_use()
    assert len(results) == len(reqs)
    for k in range(len(reqs)):
        assert results[k] == alone[k], k
    assert now == 0 and 0 < peak <= budget
    print(f"poisson: {len(reqs)} requests, at most {peak_pending[0]} submitted and not yet back, "
          f"peak device bytes {peak}", flush=True)
