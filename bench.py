"""Benchmark: decoded GB/s of one 4 GiB single-member gzip stream on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size-mib 4096]

Workload: 4 GiB (4,294,967,296 B) of synthetic log text interleaved with seeded uniform random
runs of 64 KiB-4 MiB (20% of the bytes), one gzip member, zlib level 6 with a 32 KiB window, one
stream from one compressor (no flushes; ISIZE wraps to 0; --chunk-mib N compresses in parallel
N MiB chunks ended by sync flushes instead). The compressed stream is kept in $TMPDIR for the
next run on the same machine. A step is one decode of the whole member: the compressed body is resident in
device memory before timing starts, the output goes to device memory, and the GPU's CRC-32 and
the size are checked against the member's trailer every step. After the timed steps the whole
output is compared byte for byte with the generated original.

Prints one JSON line. `roofline` is the path's: algorithmic bytes C + U per step (compressed
bytes read once, decoded bytes written once) over ms_per_step, against 8 TB/s; its `traffic` is
the HBM bytes per step of all kernels, measured on this workload (rocprofv3 FETCH_SIZE and
WRITE_SIZE passes scaled by a 1 GiB device copy, profiles/hbm_traffic_4gib_mixed_l6_one_stream.json).
`roofline.kernel` gives the same for the dominant kernel alone (most GPU time per step, from HIP
events recorded around every launch on the library's stream), per launch, with that kernel's own
algorithmic bytes. `cpu_baseline` is zlib inflate on one host thread over a 1 GiB sample of the
same content, with the speculative decoder on every host core and the reference's own sequential
decoder beside it.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_4gib_mixed_l6_one_stream.json")


def traffic_profile(C, U, per_decode=False):
    """the per-kernel HBM traffic measured on this workload (rocprofv3 FETCH_SIZE and WRITE_SIZE
    passes, each scaled by its reading for a 1 GiB device copy): {kernel: bytes per launch} (or,
    per_decode, bytes per decode of the whole stream: kernels launched inside a graph are not
    counted one by one at run time), or None when the profile is of another workload"""
    try:
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    except OSError:
        return None
    w = prof["workload"]
    if w["decoded_bytes"] != U or w["compressed_bytes"] != C:
        return None
    if per_decode:
        n = w["decodes_profiled"]
        return {k: (v["fetch_bytes_per_launch"] + v["write_bytes_per_launch"]) * v["launches"] / n
                for k, v in prof["kernels"].items()}
    return {k: v["fetch_bytes_per_launch"] + v["write_bytes_per_launch"] for k, v in prof["kernels"].items()}


def kernel_alg_bytes(kernel, C, U):
    """a kernel's algorithmic HBM bytes per decode of the whole stream (what it must move at least):
    the finder reads the input once; the decoders read it once and write 2 B per symbol; pass A
    reads and writes 2 B per byte; pass D reads 2 B and writes 1 B per byte"""
    return {"k_find": C, "k_verify": C, "k_groups": C + 2 * U, "k_decode": C + 2 * U,
            "k_tile_a": 4 * U, "k_tile_d": 3 * U, "k_stored": 3 * U}.get(kernel)


def one_stream(content, level):
    """content as one raw Deflate stream (zlib, 32 KiB window), fed to one compressor 256 MiB at a
    time (without flushes, zlib's output does not depend on how its input is split)"""
    import zlib

    import numpy as np
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    parts, step = [], 256 << 20
    for a in range(0, content.size, step):
        parts.append(c.compress(content[a:a + step]))
        log(f"compressing: {min(content.size, a + step) >> 20} of {content.size >> 20} MiB")
    parts.append(c.flush())
    return np.frombuffer(b"".join(parts), np.uint8)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(content, level, chunk):
    """CPU decoders on this host, on bounded samples of the same content compressed the same way.
    Primary (the product path for one member in the reference, src/gz.rs:132-151): zlib inflate on
    one thread. Beside it: the speculative parallel decoder (this project's planner and decoding
    core built for the host, its data-parallel steps on every core: the reference's
    research/gpu-inflate cpu_speculative in this design's terms), and the reference's own
    sequential decoder (gpuinflate/inflate.cuh built from its sources into oracle/_ref) on one core."""
    import ctypes
    import zlib

    import numpy as np

    from gpuzip_amd import datagen
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    nproc = os.cpu_count() or 1
    # (A) zlib, one thread, 1 GiB
    n = min(content.size, 1 << 30)
    s = datagen.deflate(content[:n], level, 0, chunk=chunk).tobytes()
    t = time.perf_counter()
    out = zlib.decompressobj(-15).decompress(s)
    dt = time.perf_counter() - t
    ok = len(out) == n and out == content[:n].tobytes()
    del out
    line = {"value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"zlib {zlib.ZLIB_RUNTIME_VERSION} inflate, one thread, first {n >> 20} MiB of the same content "
                      f"(zlib level {level}, {len(s)} B), {dt:.2f} s, output {'bit-exact' if ok else 'WRONG'}",
            "cpu_model": cpu_model(), "nproc": nproc}
    # (B) the speculative decoder on every core, 256 MiB
    try:
        h = oracle_lib.hostcheck()
        h.hc_set_threads.restype = None
        h.hc_set_threads.argtypes = [ctypes.c_uint]
        m = min(content.size, 256 << 20)
        sb = datagen.deflate(content[:m], level, 0, chunk=chunk).tobytes()
        h.hc_set_threads(nproc)
        t = time.perf_counter()
        rc, got, used, *_ = oracle_lib.hc_inflate(sb, m, 256 << 20, 0)
        dt = time.perf_counter() - t
        h.hc_set_threads(1)
        ok = rc == 0 and got == content[:m].tobytes()
        line["speculative_all_cores"] = {
            "value": round(m / dt / 1e9, 4), "unit": "GB/s", "cores": nproc, "kind": "port",
            "sample": f"first {m >> 20} MiB; this project's planner and decoding core built for the host, finder, "
                      f"probes, decode jobs and resolve on {nproc} threads; every bit position tested for a header "
                      f"(the reference's CPU variant stops at the first per 4 KiB chunk); {dt:.2f} s, output "
                      f"{'bit-exact' if ok else 'WRONG'}"}
    except (OSError, FileNotFoundError) as e:
        line["speculative_all_cores"] = {"error": str(e)}
    # the reference's own decoder, one core, 1 GiB
    lib = oracle_lib.ref()
    if lib is not None:
        o = np.empty(n, np.uint8)
        ol, iu, tb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
        t = time.perf_counter()
        rc = lib.ref_inflate(s, len(s), o.ctypes.data, n, ctypes.byref(ol), ctypes.byref(iu), ctypes.byref(tb))
        dt = time.perf_counter() - t
        ok = rc == 0 and ol.value == n and np.array_equal(o, content[:n])
        line["reference_decoder_one_core"] = {
            "value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "reference",
            "sample": f"gi::Decoder (gpuinflate/inflate.cuh) sequentially over the same {n >> 20} MiB, {dt:.2f} s, "
                      f"output {'bit-exact' if ok else 'WRONG'}"}
    return line


def upload_streams(streams):
    """the streams in one device buffer (each 256-byte aligned), and room for their outputs in
    another; returns (d_in, d_out, input offsets, output offsets)"""
    import numpy as np

    from gpuzip_amd.hipmem import DeviceBuffer
    offs, pos = [], 0
    for s, n, _ in streams:
        offs.append(pos)
        pos += (len(s) + 64 + 255) // 256 * 256
    host = np.zeros(pos, np.uint8)
    for (s, n, _), o in zip(streams, offs):
        host[o:o + len(s)] = np.frombuffer(s, np.uint8)
    d_in = DeviceBuffer.from_host(host)
    del host
    ooffs, opos = [], 0
    for s, n, _ in streams:
        ooffs.append(opos)
        opos += (n + 255) // 256 * 256 + 256
    return d_in, DeviceBuffer(opos), offs, ooffs


def cpu_streams_baseline(streams, seconds=10.0):
    """the same streams inflated by one zlib thread per core (the reference's ZIP CPU workers,
    src/extract.rs:416-456), over as many streams as fit in about `seconds`"""
    import zlib
    from concurrent.futures import ThreadPoolExecutor
    threads = os.cpu_count() or 8
    order = list(range(len(streams)))
    t0 = time.perf_counter()
    done, nbytes, k = 0, 0, 0
    with ThreadPoolExecutor(threads) as ex:
        while k < len(order) and time.perf_counter() - t0 < seconds:
            chunk = order[k:k + 4 * threads]
            k += len(chunk)
            outs = list(ex.map(lambda i: zlib.decompress(streams[i][0], -15), chunk))
            for i, o in zip(chunk, outs):
                if len(o) != streams[i][1] or zlib.crc32(o) != streams[i][2]:
                    raise SystemExit(f"CPU baseline: stream {i} decoded wrong")
                nbytes += len(o)
            done += len(chunk)
    dt = time.perf_counter() - t0
    return {"value": round(nbytes / dt / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"first {done} of {len(streams)} streams ({nbytes} B), zlib {zlib.ZLIB_VERSION} inflate on "
                      f"{threads} threads, {dt:.1f} s"}


def bench_batch(args):
    """Batched decode (one gi_batch_run): `streams` jobs of raw Deflate, sizes log-uniform on
    [1 KiB, 64 MiB], levels 1/6/9, inputs and outputs resident in device memory. The jobs cycle
    over `distinct` payloads (each decoded into its own output buffer); against one call per
    stream on one context, timed over the distinct payloads (the same size distribution)."""
    import numpy as np

    from gpuzip_amd import gpuinflate as gi
    from gpuzip_amd import workloads
    from gpuzip_amd.hipmem import DeviceBuffer, hip
    t = time.perf_counter()
    nd = min(args.distinct or 2000, args.streams)
    sizes = workloads.batch_sizes(nd)
    streams = workloads.make_streams(sizes, workloads.content_pool(256 << 20))
    pick = [k % nd for k in range(args.streams)]
    C = int(sum(len(streams[i][0]) for i in pick))
    U = int(sum(streams[i][1] for i in pick))
    Ud = int(sum(n for _, n, _ in streams))
    log(f"{args.streams} jobs over {nd} payloads: {U} B decoded / {C} B compressed, made in "
        f"{time.perf_counter() - t:.1f} s")
    d_in, d_out1, offs, ooffs = upload_streams(streams)
    # the schedule: chosen by the autotuner on the distinct payloads (untimed), before the other
    # output buffers take their memory
    workers, seg_mb, tuned = args.workers, args.worker_seg_mb, None
    tuning = json.loads(args.tuning) if args.tuning else None
    if not args.no_tune:
        from gpuzip_amd import autotune
        sample_jobs = (gi.Job * nd)()
        for i in range(nd):
            s, n, _ = streams[i]
            sample_jobs[i].d_in, sample_jobs[i].in_size = d_in.addr + offs[i], len(s)
            sample_jobs[i].d_out, sample_jobs[i].out_cap = d_out1.addr + ooffs[i], n + 256
        expected = [(n, len(s), crc) for s, n, crc in streams]
        with gi.Gpu(device=0, segment_out_mb=64) as g0:  # a small context, for the device's name
            key = f"batch:{g0.device}:{nd}:{sum(sizes)}"
        tuned = autotune.tune_batch(sample_jobs, expected, cache=args.tune_cache, key=key, log=log)
        p = tuned["params"]
        workers, seg_mb, tuning = p["workers"], p["segment_out_mb"], autotune.tuning_of(p)
        log(f"autotuner: {p} at {tuned['GBs']} GB/s on the {nd} payloads (defaults {tuned.get('default_GBs')}), "
            f"{tuned['trials']} trials in {tuned['seconds']} s{' (cached)' if tuned['cached'] else ''}")
        del sample_jobs
    # every job its own output: one buffer per pass over the payloads
    reps = (args.streams + nd - 1) // nd
    outs = [d_out1] + [DeviceBuffer(d_out1.nbytes) for _ in range(reps - 1)]
    jobs = (gi.Job * args.streams)()
    for k, i in enumerate(pick):
        s, n, _ = streams[i]
        jobs[k].d_in, jobs[k].in_size = d_in.addr + offs[i], len(s)
        jobs[k].d_out, jobs[k].out_cap = outs[k // nd].addr + ooffs[i], n + 256

    def check(tag, sample=()):
        for k, i in enumerate(pick):
            s, n, crc = streams[i]
            j = jobs[k]
            if j.status != gi.GI_OK or j.out_size != n or j.crc32 != crc or j.in_used != len(s):
                raise SystemExit(f"{tag}: job {k} failed: status {j.status} size {j.out_size}/{n}")
        import zlib
        for k in sample:  # the bytes themselves, where they were meant to go
            i = pick[k]
            got = outs[k // nd].to_host(streams[i][1], offset=ooffs[i]).tobytes()
            if got != zlib.decompress(streams[i][0], -15):
                raise SystemExit(f"{tag}: job {k}: output bytes differ")

    rng = np.random.default_rng(7)
    sample = sorted(set(rng.integers(0, args.streams, 200).tolist()) | {int(np.argmax([streams[i][1] for i in pick]))})
    # one call per stream, one context, over the distinct payloads
    g = gi.Gpu(device=0)
    hip().hipDeviceSynchronize()
    for k in range(min(8, nd)):
        g.inflate_device(jobs[k].d_in, jobs[k].in_size, jobs[k].d_out, jobs[k].out_cap)
    t0 = time.perf_counter()
    for k in range(nd):
        j = jobs[k]
        osz, c, st = g.inflate_device(j.d_in, j.in_size, j.d_out, j.out_cap)
        j.status, j.out_size, j.crc32, j.in_used = 0, osz, c, st.in_used
        if k % 500 == 499:
            log(f"  per-stream calls: {k + 1} done")
    t_single = time.perf_counter() - t0
    single_gbs = Ud / t_single / 1e9
    log(f"per-stream calls: {nd} streams in {t_single:.2f} s, {single_gbs:.2f} GB/s")
    g.close()
    for k in range(args.streams):
        jobs[k].status = -1
    with gi.Batch(workers=workers, segment_out_mb=seg_mb, tuning=tuning) as b:
        t0 = time.perf_counter()
        b.run(jobs)  # warm-up
        log(f"batch warm-up: {time.perf_counter() - t0:.2f} s")
        check("batch warm-up", sample)
        b.kernel_times(reset=True)
        times = []
        for step in range(args.steps):
            for k in range(args.streams):
                jobs[k].status = -1
            t0 = time.perf_counter()
            b.run(jobs)
            times.append(time.perf_counter() - t0)
            check("batch")
            log(f"batch step {step}: {times[-1]:.2f} s")
        kt, nseg, nalone = b.kernel_times()
    elapsed = sum(times)
    # launches per batch: each batch segment is a find graph, one group-decoder launch a decode
    # round (and a lane-decoder launch where a round has lane jobs), and a resolve graph, for all
    # of its streams at once
    launches = {k: {"launches": round(v[0] / args.steps, 1), "gpu_ms": round(v[1] / args.steps * 1e3, 2)}
                for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])}
    # (gpu_ms: HIP-event time summed over the workers, whose launches overlap)
    per_batch = {"batch_segments": round(nseg / args.steps, 1), "streams_alone": round(nalone / args.steps, 1),
                 "launches": launches, "launches_total": round(sum(v[0] for v in kt.values()) / args.steps, 1),
                 "streams": args.streams}
    log(f"per batch: {per_batch}")
    value = U * args.steps / elapsed / 1e9
    log(f"batched: {elapsed / args.steps:.2f} s per batch, {value:.2f} GB/s ({workers} workers)")
    cpu = None if args.no_cpu else cpu_streams_baseline(streams)
    line = {
        "metric": "batched decoded GB/s, raw-deflate streams log-uniform 1 KiB-64 MiB, levels 1/6/9",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{args.streams} raw Deflate streams in one gi_batch_run, sizes log-uniform on "
                               f"[1 KiB, 64 MiB], zlib levels 1/6/9 in equal shares, log text with 20% random runs; "
                               f"{nd} distinct payloads, each decoded into its own output buffer",
                   "streams": args.streams, "distinct_payloads": nd, "decoded_bytes": U, "compressed_bytes": C,
                   "device_resident": True, "workers": workers, "worker_segment_mb": seg_mb,
                   "schedule": dict(tuning or {}, source="autotuner" if tuned else "defaults"),
                   "crc_checked": True, "bytes_checked_jobs": len(sample)},
        "autotune": tuned,
        "per_batch": per_batch,
        # the path's roofline: compressed bytes read once and decoded bytes written once, per batch
        "roofline": {"bound": "hbm", "achieved": round((C + U) / (elapsed / args.steps) / 1e9, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round((C + U) / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 5),
                     "traffic": None, "alg_bytes_per_step": C + U,
                     "scope": "path: (C + U) per batch over the batch's time; traffic not profiled for this workload"},
        "per_stream_calls_GBs": round(single_gbs, 3),
        "batched_over_per_stream": round(value / single_gbs, 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_rescue_headers(args):
    """The rescue's first step (gpuinflate/rescue.cpp:597-626, every confirmed dynamic-block header
    of a damaged file) on the GPU: a gzip member of `size-mib` MiB (default 1 GiB) of the
    single-stream content at level 6, three 512-byte sectors zeroed at seeded offsets. Timed from
    host memory to the host list; the CPU baseline is the reference's own rescue.cpp (oracle/_ref,
    16 threads) on the file's first 64 MiB, where the two lists are compared too."""
    import numpy as np

    from gpuzip_amd import datagen, gz
    from gpuzip_amd.gpuinflate import Gpu
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    size = (args.size_mib if args.size_mib != 4096 else 1024) << 20
    t = time.perf_counter()
    content = datagen.mixed(size, seed=0xC2)
    body = datagen.deflate(content, args.level, 0, chunk=args.chunk_mib << 20)
    member = bytearray(datagen.gzip_member(body, size, datagen.crc32(content)))
    del content, body
    rng = np.random.default_rng(0xD2)
    nsec = len(member) // 512
    sectors = sorted(rng.choice(np.arange(1, nsec - 1), 3, replace=False).tolist())
    for sct in sectors:
        member[sct * 512:(sct + 1) * 512] = bytes(512)
    data = bytes(member)
    del member
    hlen = gz.header(data[:4096])[0]
    log(f"{len(data)} B gzip ({size} B decoded), sectors {sectors} zeroed, made in {time.perf_counter() - t:.1f} s")
    g = Gpu(device=0)
    g.rescue_headers(data[:1 << 20], hlen * 8)  # warm-up
    times, hs = [], None
    for _ in range(args.steps):
        t = time.perf_counter()
        hs = g.rescue_headers(data, hlen * 8)
        times.append(time.perf_counter() - t)
    ms = float(np.median(times)) * 1e3
    value = len(data) / (ms / 1e3) / 1e9
    log(f"GPU: {len(hs)} headers, median {ms:.1f} ms, {value:.2f} GB/s of compressed input")
    cpu, parity = None, None
    os.environ.setdefault("GR_THREADS", "16")
    from oracle_lib import ref_rescue_headers
    sample = data[:64 << 20]
    t = time.perf_counter()
    want = ref_rescue_headers(sample, hlen * 8)
    dt = time.perf_counter() - t
    if want is not None:
        got = g.rescue_headers(sample, hlen * 8)
        parity = got == want
        if not parity:
            raise SystemExit("rescue headers differ from the reference's on the sample")
        cpu = {"value": round(len(sample) / dt / 1e9, 4), "unit": "GB/s", "cores": int(os.environ["GR_THREADS"]),
               "kind": "reference", "sample": f"the file's first {len(sample)} B, gpuinflate/rescue.cpp find_headers "
                                              f"built from its sources, {dt:.1f} s, {len(want)} headers (equal to the GPU's)"}
    line = {
        "metric": "rescue header search GB/s (compressed input), damaged 1 GiB-decoded gzip",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"gzip member of {size} B of log text with 20% random runs, zlib level {args.level}, "
                               f"three 512 B sectors zeroed at seeded offsets; every confirmed dynamic header",
                   "compressed_bytes": len(data), "headers": len(hs), "host_input": True},
        "parity_with_reference_sample": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


RESCUE_KERNEL_BYTES = {
    # algorithmic HBM bytes per step of each rescue kernel, from C (compressed bytes), U (decoded
    # bytes) and S (bytes the decoders copied from stored blocks): the decoders read the input once
    # and write a 16-bit symbol a byte; the stored copies read a byte and write a symbol; the
    # provenance set-up reads a symbol and writes a 64-bit entry a byte; a jump pass reads every
    # entry; the split reads an entry and writes the byte and its 32-bit tag
    "k_rescue_decode": lambda C, U, S: C + 2 * (U - S),
    "k_rescue_stored": lambda C, U, S: 3 * S,
    "k_prov_init": lambda C, U, S: 10 * U,
    "k_prov_jump": lambda C, U, S: 8 * U,
    "k_prov_split": lambda C, U, S: 13 * U,
}


def bench_rescue(args):
    """Rescue of a damaged file on the GPU (libgzrescue.so through gpuzip_amd.rescue, the
    reference loader's interface): a gzip member of `size-mib` MiB (default 1 GiB) of the
    single-stream content (log text with 20% random runs), one zlib call at level 6, three 512-byte
    sectors zeroed at seeded offsets. A step is one whole rescue, host memory to the host sink
    (the interface's own boundary: PCIe both ways included). The reference's rescue
    (gpuinflate/rescue.cpp built from its sources, oracle/_ref) on the same file and the host's
    cores is the CPU baseline, and its output, mask, statistics, pieces and notes are compared
    with the GPU's."""
    import hashlib

    import numpy as np

    from gpuzip_amd import datagen
    from gpuzip_amd import rescue as gr
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    size = (args.size_mib if args.size_mib != 4096 else 1024) << 20
    t = time.perf_counter()
    # one zlib call over 1 GiB takes the better part of a minute: the file is kept for the next
    # run on the same machine (same size and level give the same bytes)
    cache = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gpuzip_rescue_{size}_l{args.level}.gz")
    if os.path.exists(cache):
        member = bytearray(open(cache, "rb").read())
    else:
        content = datagen.mixed(size)
        body = datagen.deflate(content, args.level, 0)
        member = bytearray(datagen.gzip_member(body, size, datagen.crc32(content)))
        del content, body
        with open(cache + ".part", "wb") as f:
            f.write(member)
        os.replace(cache + ".part", cache)
    C = len(member) - 18
    rng = np.random.default_rng(0xD2)
    nsec = len(member) // 512
    sectors = sorted(rng.choice(np.arange(1, nsec - 1), 3, replace=False).tolist())
    for sct in sectors:
        member[sct * 512:(sct + 1) * 512] = bytes(512)
    data = bytes(member)
    del member
    log(f"{len(data)} B gzip ({size} B decoded, one zlib call), sectors {sectors} zeroed, ready in "
        f"{time.perf_counter() - t:.1f} s")
    rs = gr.get()
    rs.rescue(data[:1 << 20], 0, lambda d, m: True)  # warm-up: device, header search context
    got = {"out": 0}

    def count(d, m):
        got["out"] += len(d)
        return True

    times, rep = [], None
    for k in range(args.steps):
        got["out"] = 0
        t = time.perf_counter()
        rep = rs.rescue(data, 0, count)
        times.append(time.perf_counter() - t)
        log(f"step {k}: {times[-1]:.3f} s, {got['out']} B out; " +
            ", ".join(f"{n} {v:.3f}" for n, v in rep.phases.items()))
    U = got["out"]
    s_med = float(np.median(times))
    value = U / s_med / 1e9
    for name, (launches, sec) in sorted(rep.kernels.items(), key=lambda kv: -kv[1][1]):
        log(f"  {name:16s} {launches:5d} launches {sec * 1e3:9.2f} ms")
    # parity: the GPU's output and mask hashed as they arrive, against the reference's
    h_out, h_mask = hashlib.sha256(), hashlib.sha256()

    def hashed(d, m):
        h_out.update(d)
        h_mask.update(m)
        return True

    rep = rs.rescue(data, 0, hashed)
    parity, cpu = None, None
    if not args.no_cpu:
        import rescue_lib
        ref = rescue_lib.reference()
        if ref is not None:
            os.environ.setdefault("GR_THREADS", "16")
            t = time.perf_counter()
            want = rescue_lib.rescue(ref, data, max_output=gr.default_max_output(len(data)))
            dt = time.perf_counter() - t
            pieces = [(p.out_offset, p.length, p.decoded, p.known, p.in_byte_start, p.in_byte_end, int(p.suspect))
                      for p in rep.pieces]
            parity = (want[0] == 0 and hashlib.sha256(want[1]).hexdigest() == h_out.hexdigest() and
                      hashlib.sha256(want[2]).hexdigest() == h_mask.hexdigest() and want[3] == rep.stats and
                      [w[:7] for w in want[4]] == pieces and want[5] == rep.notes)
            log(f"reference rescue: {dt:.1f} s on {os.environ['GR_THREADS']} threads; parity {parity}")
            if not parity:
                raise SystemExit("the GPU rescue differs from the reference's")
            cpu = {"value": round(len(want[1]) / dt / 1e9, 4), "unit": "GB/s", "cores": int(os.environ["GR_THREADS"]),
                   "kind": "reference", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                   "sample": f"the whole file: gpuinflate/rescue.cpp's gr_rescue built from its sources, "
                             f"{os.environ['GR_THREADS']} threads, {dt:.1f} s; output, mask, stats, pieces and notes "
                             f"equal to the GPU's"}
    stored = rep.stored_bytes
    kern = {}
    for name, (launches, sec) in rep.kernels.items():
        kern[name] = {"launches": launches, "ms": round(sec * 1e3, 3)}
    dom = max(rep.kernels.items(), key=lambda kv: kv[1][1])
    dname, (dl, dsec) = dom
    k_alg = RESCUE_KERNEL_BYTES.get(dname, lambda C, U, S: None)(C, U, stored)
    k_ach = k_alg / dsec / 1e9 if k_alg else None
    alg = C + U
    achieved = alg / s_med / 1e9
    line = {
        "metric": "rescue GB/s (decoded), damaged 1 GiB-decoded gzip, host to host",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(s_med * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"gzip member of {size} B of log text with 20% random runs of 64 KiB-4 MiB, one zlib "
                               f"call at level {args.level}, 32 KiB window, three 512 B sectors zeroed at seeded "
                               f"offsets {sectors}", "compressed_bytes": len(data), "decoded_bytes": U,
                   "headers": rep.stats["headers"], "pieces": rep.stats["pieces"], "damaged": rep.stats["damaged"],
                   "host_input_and_output": True},
        "parity_with_reference": parity,
        "phases_s": {k: round(v, 4) for k, v in rep.phases.items()},
        "kernels": kern,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "alg_bytes_per_step": alg,
                     "scope": "path: (C + U) per rescue over ms_per_step",
                     "kernel": {"name": dname, "launches_per_step": dl, "ms_per_launch": round(dsec / dl * 1e3, 4),
                                "alg_bytes_per_launch": int(k_alg / dl) if k_alg else None,
                                "achieved": round(k_ach, 2) if k_ach else None,
                                "frac": round(k_ach / HBM_PEAK_GBS, 5) if k_ach else None}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def serve_roofline(alg_bytes, seconds):
    """a served run's roofline: the requests' compressed and decoded bytes (each read or written
    once) over the run, against the HBM peak; the load is the offered one, so this is the share
    of the chip the served rate needs, not a limit it reached"""
    a = alg_bytes / seconds / 1e9
    return {"bound": "hbm", "achieved": round(a, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(a / HBM_PEAK_GBS, 5),
            "traffic": None, "scope": "path: the served requests' C + U over the run"}


def bench_serve(args):
    """Serving: gzip members arriving as a Poisson process at a fixed offered load; each request's
    header parsed on the host (libgzhost gz_header), admitted against an HBM budget (input +
    output + the decoders' 16-bit symbols), submitted to a gi_batch (what arrives while the
    workers are busy forms their next batch segment); latency from arrival to the result checked
    against the member's trailer (CRC-32, ISIZE, where the trailer lies)"""
    import numpy as np

    from gpuzip_amd import datagen
    from gpuzip_amd import gpuinflate as gi
    from gpuzip_amd import gz
    from gpuzip_amd import workloads
    from gpuzip_amd.hipmem import DeviceBuffer
    rng = np.random.default_rng(0x6E)
    ndist = args.distinct or 512
    sizes = workloads.serve_sizes(ndist)
    raws = workloads.make_streams(sizes, workloads.content_pool(256 << 20))
    # gzip members: (member bytes, output size, CRC-32)
    streams = [(datagen.gzip_member(r, n, c), n, c) for r, n, c in raws]
    del raws
    mean_u = float(np.mean([n for _, n, _ in streams]))
    lam = args.offered * 1e9 / mean_u
    arrivals = np.cumsum(rng.exponential(1 / lam, int(lam * args.seconds * 1.2) + 16))
    arrivals = arrivals[arrivals < args.seconds]
    pick = rng.integers(0, len(streams), arrivals.size)
    log(f"{arrivals.size} requests over {args.seconds} s (offered {args.offered} GB/s, mean {mean_u / 2**20:.2f} MiB)")
    d_in, _, offs, _ = upload_streams(streams)
    budget = args.budget_gib << 30
    arena = DeviceBuffer(min(budget // 2, 48 << 30))
    free = [(0, arena.nbytes)]  # first fit over the output arena

    def alloc(n):
        n = (n + 4095) // 4096 * 4096
        for i, (a, sz) in enumerate(free):
            if sz >= n:
                free[i] = (a + n, sz - n)
                return a
        return None

    def release(a, n):
        n = (n + 4095) // 4096 * 4096
        free.append((a, n))
        free.sort()
        merged = []
        for x, sz in free:
            if merged and merged[-1][0] + merged[-1][1] == x:
                merged[-1] = (merged[-1][0], merged[-1][1] + sz)
            elif sz:
                merged.append((x, sz))
        free[:] = merged

    jobs = (gi.Job * arrivals.size)()
    hlen = np.zeros(arrivals.size, np.int64)
    lat, done_bytes, inflight_bytes, done_c = [], 0, 0, 0
    pending, regions = [], {}
    with gi.Batch(workers=args.workers, segment_out_mb=args.worker_seg_mb) as b:
        t0 = time.perf_counter()
        nxt, finished = 0, 0
        while finished < arrivals.size:
            now = time.perf_counter() - t0
            while nxt < arrivals.size and arrivals[nxt] <= now:
                pending.append(nxt)
                nxt += 1
            # admission: in arrival order, while the budget and the arena allow
            while pending:
                k = pending[0]
                s, n, _ = streams[pick[k]]
                cost = len(s) + 5 * n  # input, output, the decoders' 16-bit symbols
                if inflight_bytes + cost > budget:
                    break
                a = alloc(n + 256)
                if a is None:
                    break
                pending.pop(0)
                h = int(gz.header(s[:4096])[0])  # the member's header, parsed on the host
                hlen[k] = h
                j = jobs[k]
                j.d_in, j.in_size = d_in.addr + offs[pick[k]] + h, len(s) - h
                j.d_out, j.out_cap = arena.addr + a, n + 256
                regions[k] = (a, n + 256, cost)
                inflight_bytes += cost
                b.submit(j)
            wait = 0.0005 if (pending or nxt < arrivals.size) else 1.0
            addr = b.wait(wait)
            while addr:
                k = (addr - ctypes.addressof(jobs)) // ctypes.sizeof(gi.Job)
                j = jobs[k]
                s, n, _ = streams[pick[k]]
                t = int(hlen[k]) + j.in_used  # the trailer: CRC-32 and ISIZE
                if (j.status != gi.GI_OK or t + 8 != len(s) or j.crc32 != int.from_bytes(s[t:t + 4], "little")
                        or j.out_size & 0xFFFFFFFF != int.from_bytes(s[t + 4:t + 8], "little")):
                    raise SystemExit(f"request {k} failed: status {j.status}")
                lat.append(time.perf_counter() - t0 - arrivals[k])
                done_bytes += n
                done_c += len(s)
                a, sz, cost = regions.pop(k)
                release(a, sz)
                inflight_bytes -= cost
                finished += 1
                if finished % 2000 == 0:
                    log(f"  {finished} of {arrivals.size} done")
                addr = b.wait(0.0)
        elapsed = time.perf_counter() - t0
    lat = np.array(lat) * 1e3
    value = done_bytes / elapsed / 1e9
    log(f"served {arrivals.size} requests in {elapsed:.2f} s: {value:.2f} GB/s, p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms")
    line = {
        "metric": f"served decoded GB/s with p50/p99 latency, Poisson arrivals at {args.offered:g} GB/s offered",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(elapsed * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{arrivals.size} gzip-member requests over {args.seconds} s, sizes lognormal median "
                               f"2 MiB sigma 1.5 capped at 512 MiB ({ndist} distinct payloads), levels 1/6/9, "
                               f"log text with 20% random runs",
                   "offered_GBs": args.offered, "hbm_budget_gib": args.budget_gib, "workers": args.workers,
                   "device_resident": True, "crc_checked": True},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "roofline": serve_roofline(done_c + done_bytes, elapsed),
    }
    print(json.dumps(line), flush=True)


def bench_serve_library(args):
    """Serving through the library's own server (gi_serve_*): the same Poisson trace of gzip members
    as bench_serve, but each request is a gzip member in host memory with a host output buffer.
    The server parses its header, admits it in arrival order when its input and output fit the
    HBM budget, uploads it, folds it into the batch workers' next segment, copies the output back
    and checks the member's trailer (CRC-32 from the GPU, ISIZE). Latency is from arrival to the
    output in host memory, so it includes both PCIe copies; inputs and outputs live in pinned host
    memory, as a server's network buffers would."""
    import queue
    import threading
    import zlib

    import numpy as np

    from gpuzip_amd import datagen
    from gpuzip_amd import gpuinflate as gi
    from gpuzip_amd import workloads
    from gpuzip_amd.hipmem import check, hip
    rng = np.random.default_rng(0x6E)
    ndist = args.distinct or 512
    sizes = workloads.serve_sizes(ndist)
    raws = workloads.make_streams(sizes, workloads.content_pool(256 << 20))
    streams = [(datagen.gzip_member(r, n, c), n, c) for r, n, c in raws]
    del raws
    mean_u = float(np.mean([n for _, n, _ in streams]))
    lam = args.offered * 1e9 / mean_u
    arrivals = np.cumsum(rng.exponential(1 / lam, int(lam * args.seconds * 1.2) + 16))
    arrivals = arrivals[arrivals < args.seconds]
    pick = rng.integers(0, len(streams), arrivals.size)
    log(f"library server: {arrivals.size} requests over {args.seconds} s (offered {args.offered} GB/s, "
        f"mean {mean_u / 2**20:.2f} MiB)")
    h = hip()
    h.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
    h.hipHostFree.argtypes = [ctypes.c_void_p]

    def pinned(n):
        p = ctypes.c_void_p()
        check(h.hipHostMalloc(ctypes.byref(p), max(n, 1), 0), "hipHostMalloc")
        return p.value

    # the distinct members, once, in pinned memory
    offs, pos = [], 0
    for s, _, _ in streams:
        offs.append(pos)
        pos += (len(s) + 255) // 256 * 256
    in_base = pinned(pos)
    for (s, _, _), o in zip(streams, offs):
        ctypes.memmove(in_base + o, s, len(s))
    # output room, the caller's buffers: four in pinned memory for each distinct payload (a request
    # takes a free one of its payload's, and waits for one when all are in flight, as a server
    # waits for a client's buffer); each payload comes back about every 150 ms at this load
    COPIES = 4
    out_offs, pos = [], 0
    for _, n, _ in streams:
        out_offs.append(pos)
        pos += COPIES * ((n + 4095) // 4096 * 4096)
    out_base = pinned(pos)
    busy = [[False] * COPIES for _ in streams]
    lock = threading.Lock()

    def alloc(i):
        for c in range(COPIES):
            if not busy[i][c]:
                busy[i][c] = True
                return out_offs[i] + c * ((streams[i][1] + 4095) // 4096 * 4096), c
        return None

    reqs = (gi.Request * arrivals.size)()
    region = {}
    lat, failed = np.zeros(arrivals.size), []
    adm, srv_s = np.zeros(arrivals.size), np.zeros(arrivals.size)  # admission wait; submission to output
    lag = np.zeros(arrivals.size)  # from arrival to submission (a client behind its arrivals)
    done = [0, 0, 0]  # requests, decoded bytes, compressed bytes
    srv = gi.Server(args.budget_gib << 30, workers=args.workers, segment_out_mb=args.worker_seg_mb)
    t0 = time.perf_counter()

    def completer():
        received = 0
        while received < arrivals.size:
            p = srv._lib.gi_serve_wait(srv._s, 1.0)
            if not p:
                continue
            received += 1
            k = (ctypes.addressof(p.contents) - ctypes.addressof(reqs)) // ctypes.sizeof(gi.Request)
            r = reqs[k]
            _, n, crc = streams[pick[k]]
            lat[k] = time.perf_counter() - t0 - arrivals[k]
            adm[k], srv_s[k] = r.admitted_s, r.done_s
            if r.status != gi.GI_OK or r.out_size != n or r.crc32 != crc:
                failed.append((k, r.status, r.error.decode(errors="replace")))
            elif k % 256 == 0:
                # the bytes delivered, sampled: checked by another thread, the buffer held until
                # then, so that the check does not delay the completions after this one
                checks.put(k)
                continue
            finish(k, n)

    def finish(k, n):
        with lock:
            busy[pick[k]][region.pop(k)] = False
            done[0] += 1
            done[1] += n
            done[2] += len(streams[pick[k]][0])
        if done[0] % 4000 == 0:
            log(f"  {done[0]} of {arrivals.size} done")

    checks = queue.Queue()

    def checker():
        while True:
            k = checks.get()
            if k is None:
                return
            _, n, crc = streams[pick[k]]
            if zlib.crc32((ctypes.c_char * n).from_address(reqs[k].out)) != crc:
                failed.append((k, reqs[k].status, "delivered bytes differ"))
            finish(k, n)

    th = threading.Thread(target=completer)
    th.start()
    ck = threading.Thread(target=checker)
    ck.start()
    try:
        for k in range(arrivals.size):
            s, n, _ = streams[pick[k]]
            while True:
                with lock:
                    got = alloc(pick[k])
                    if got is not None:
                        a, region[k] = got
                        break
                time.sleep(0.0002)
            wait = arrivals[k] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            r = reqs[k]
            r.in_ = in_base + offs[pick[k]]
            r.in_size = len(s)
            r.out = out_base + a
            r.out_cap = n
            r.gzip = 1
            lag[k] = time.perf_counter() - t0 - arrivals[k]
            rc = srv._lib.gi_serve_submit(srv._s, ctypes.byref(r))
            if rc != gi.GI_OK:
                raise SystemExit(f"submit {k}: {rc}")
        th.join()
    finally:
        elapsed = time.perf_counter() - t0
        checks.put(None)
        ck.join()
        srv.close()
    assert done[0] == arrivals.size or failed
    if failed:
        raise SystemExit(f"{len(failed)} requests failed, the first: {failed[0]}")
    value = done[1] / elapsed / 1e9
    lat *= 1e3
    log(f"served {arrivals.size} requests in {elapsed:.2f} s: {value:.2f} GB/s, p50 {np.percentile(lat, 50):.1f} ms, "
        f"p99 {np.percentile(lat, 99):.1f} ms")
    line = {
        "metric": f"served decoded GB/s with p50/p99 latency, Poisson arrivals at {args.offered:g} GB/s offered, "
                  f"library server (host memory in and out)",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(elapsed * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{arrivals.size} gzip-member requests over {args.seconds} s, sizes lognormal median "
                               f"2 MiB sigma 1.5 capped at 512 MiB ({ndist} distinct payloads), levels 1/6/9, "
                               f"log text with 20% random runs",
                   "offered_GBs": args.offered, "hbm_budget_gib": args.budget_gib, "workers": args.workers,
                   "device_resident": False, "pcie_in_and_out": True, "trailers_checked": True},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        # inside the server: from submission to admission (device memory free), and to the output
        # in host memory
        "roofline": serve_roofline(done[2] + done[1], elapsed),
        "server_ms": {"admission_p50": round(float(np.percentile(adm, 50)) * 1e3, 3),
                      "admission_p99": round(float(np.percentile(adm, 99)) * 1e3, 3),
                      "done_p50": round(float(np.percentile(srv_s, 50)) * 1e3, 3),
                      "done_p99": round(float(np.percentile(srv_s, 99)) * 1e3, 3),
                      "client_lag_p50": round(float(np.percentile(lag, 50)) * 1e3, 3),
                      "client_lag_p99": round(float(np.percentile(lag, 99)) * 1e3, 3)},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size-mib", type=int, default=4096)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--chunk-mib", type=int, default=0,
                    help="0: one zlib stream; N: parallel N MiB chunks ended by sync flushes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", action="store_true", help="the batched workload instead of the single stream")
    ap.add_argument("--serve", action="store_true", help="the served workload instead of the single stream")
    ap.add_argument("--library", action="store_true",
                    help="--serve through the library's server (gi_serve_*), requests and outputs in host memory")
    ap.add_argument("--streams", type=int, default=20000, help="--batch: jobs in the batch")
    ap.add_argument("--workers", type=int, default=2, help="--batch/--serve: decoder contexts")
    ap.add_argument("--worker-seg-mb", type=int, default=0,
                    help="--batch/--serve: a worker's segment (0: sized to the free memory, at most 4 GiB)")
    ap.add_argument("--no-tune", action="store_true", help="--batch: the library's default schedule, no autotuner")
    ap.add_argument("--tuning", default=None,
                    help="a schedule as JSON (gi_config_tuned's fields, e.g. '{\"lanes_per_group\": 64}'): the single "
                         "stream's, or --batch's with --no-tune")
    ap.add_argument("--tune-cache", default=None, help="--batch: JSON file the autotuner's choice is kept in")
    ap.add_argument("--seconds", type=float, default=10.0, help="--serve: arrivals over this long")
    ap.add_argument("--offered", type=float, default=20.0, help="--serve: offered load, GB/s decoded")
    ap.add_argument("--distinct", type=int, default=0,
                    help="distinct payloads (--batch: 2000, --serve: 512)")
    ap.add_argument("--budget-gib", type=int, default=64, help="--serve: HBM admission budget")
    ap.add_argument("--rescue", action="store_true",
                    help="the rescue of a damaged 1 GiB gzip file (GPU against the reference's rescue) instead")
    ap.add_argument("--rescue-headers", action="store_true",
                    help="the rescue's header search on a damaged 1 GiB gzip file instead of the single stream")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one device-to-device copy of this many MiB before the steps: the known-size "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this decoder has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)
    if args.rescue:
        return bench_rescue(args)
    if args.rescue_headers:
        return bench_rescue_headers(args)
    if args.batch:
        return bench_batch(args)
    if args.serve:
        return bench_serve_library(args) if args.library else bench_serve(args)

    import numpy as np

    from gpuzip_amd import datagen
    from gpuzip_amd import gpuinflate as gi
    from gpuzip_amd.gpuinflate import Gpu
    from gpuzip_amd.hipmem import D2D, DeviceBuffer, check, hip

    U = args.size_mib << 20
    t = time.perf_counter()
    content = datagen.mixed(U)
    log(f"generated {U} B in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    # one zlib call over gigabytes takes minutes: the stream is kept for the next run on the same
    # machine (the same size, level and chunking give the same bytes)
    cache = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gpuzip_bench_{U}_l{args.level}_c{args.chunk_mib}.deflate")
    if os.path.exists(cache):
        body = np.fromfile(cache, np.uint8)
    else:
        body = one_stream(content, args.level) if args.chunk_mib == 0 else \
            datagen.deflate(content, args.level, 0, chunk=args.chunk_mib << 20)
        body.tofile(cache + ".part")
        os.replace(cache + ".part", cache)
    crc = datagen.crc32(content)
    log(f"compressed to {body.size} B (C/U {body.size / U:.3f}) in {time.perf_counter() - t:.1f} s")
    # the member: header (10 B), body, trailer (CRC-32, ISIZE = size mod 2^32)
    isize = U & 0xFFFFFFFF
    C = body.size

    tuning = json.loads(args.tuning) if args.tuning else None
    g = Gpu(device=0, tuning=tuning)
    d_in = DeviceBuffer.from_host(body)
    d_out = DeviceBuffer(U)
    hip().hipDeviceSynchronize()
    log(f"device {g.device}: input resident ({C} B), output buffer {U} B")
    if args.copy_mib:
        n = min(args.copy_mib << 20, U, C)
        check(hip().hipMemcpy(ctypes.c_void_p(d_out.addr), ctypes.c_void_p(d_in.addr), n, D2D), "hipMemcpy D2D")
        hip().hipDeviceSynchronize()
        log(f"copied {n} B device to device (counter scale)")

    def step():
        n, c, st = g.inflate_device(d_in.addr, C, d_out.addr, U, isize)
        if c != crc or (n & 0xFFFFFFFF) != isize or st.in_used != C:
            raise SystemExit(f"trailer check failed: size {n} crc {c:08x} (want {crc:08x}), in_used {st.in_used}/{C}")
        return st

    for w in range(args.warmup):
        t = time.perf_counter()
        st = step()
        log(f"warmup {w}: {time.perf_counter() - t:.3f} s, {st.segments} segments, {st.rounds} rounds, {st.jobs} jobs")
    ktimes = {}
    hip().hipDeviceSynchronize()
    t0 = time.perf_counter()
    per_step = []
    for k in range(args.steps):
        ts = time.perf_counter()
        st = step()
        per_step.append(time.perf_counter() - ts)
        for name, (launches, sec) in g.kernel_times().items():
            a = ktimes.setdefault(name, [0, 0.0])
            a[0] += launches
            a[1] += sec
    hip().hipDeviceSynchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    value = U * args.steps / elapsed / 1e9
    log(f"timed {args.steps} steps: {ms:.1f} ms/step, {value:.2f} GB/s decoded; "
        f"stages find {st.find_s * 1e3:.1f} decode {st.decode_s * 1e3:.1f} resolve {st.resolve_s * 1e3:.1f} ms")
    per_step.sort()
    log(f"  steps: min {per_step[0] * 1e3:.1f} median {per_step[len(per_step) // 2] * 1e3:.1f} "
        f"max {per_step[-1] * 1e3:.1f} ms")
    for name, (launches, sec) in sorted(ktimes.items(), key=lambda kv: -kv[1][1]):
        log(f"  {name:12s} {launches:6d} launches {sec * 1e3 / args.steps:9.2f} ms/step")

    out = d_out.to_host()
    exact = bool(np.array_equal(out, content))
    del out
    if not exact:
        raise SystemExit("output differs from the original")
    log("output bit-exact with the original")

    # the host path (gi_inflate): input from host memory, output through a sink, PCIe both ways;
    # reported beside the device-resident value, never as it
    delivered = [0]

    def count(user, ptr, n):
        delivered[0] += n
        return 0

    cb = gi.SINK(count)
    hst = gi.Stats()
    t = time.perf_counter()
    with g._lock:
        rc = g._lib.gi_inflate(g._ctx, ctypes.cast(body.ctypes.data, ctypes.c_char_p), C, isize, cb, None,
                               ctypes.byref(hst))
    host_s = time.perf_counter() - t
    if rc != gi.GI_OK or delivered[0] != U or g.last_crc32() != crc:
        raise SystemExit(f"host path failed: rc {rc}, {delivered[0]} B delivered")
    log(f"host path (gi_inflate, PCIe in and out): {host_s:.3f} s, {U / host_s / 1e9:.2f} GB/s")

    # The path: its algorithmic bytes (compressed read once, decoded written once) over the step
    # time; traffic: every kernel's measured HBM bytes per step (rocprofv3 counters on this
    # workload, a step being one decode). The dominant kernel (most GPU time per step;
    # "find_graph" and "resolve_graph" are the finder's and the resolve's kernels as one graph
    # each, not kernels) with its own algorithmic bytes, all per launch.
    alg = C + U
    achieved = alg / (ms / 1e3) / 1e9
    prof = traffic_profile(C, U)
    per_decode = traffic_profile(C, U, per_decode=True)
    traffic = int(sum(per_decode.values())) if per_decode else None
    dom, (dl, ds) = max(((k, v) for k, v in ktimes.items() if k.startswith("k_")), key=lambda kv: kv[1][1])
    k_alg = kernel_alg_bytes(dom, C, U)
    per_launch = dl / args.steps
    k_launch_s = ds / dl
    k_alg_launch = k_alg / per_launch if k_alg else None
    k_ach = k_alg_launch / k_launch_s / 1e9 if k_alg_launch else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "alg_bytes_per_step": alg,
                "scope": "path: (C + U) per step over ms_per_step; traffic: HBM bytes per step, all kernels",
                "kernel": {"name": dom, "launches_per_step": per_launch, "ms_per_launch": round(k_launch_s * 1e3, 4),
                           "alg_bytes_per_launch": int(k_alg_launch) if k_alg_launch else None,
                           "achieved": round(k_ach, 2) if k_ach else None,
                           "frac": round(k_ach / HBM_PEAK_GBS, 5) if k_ach else None,
                           "traffic_per_launch": prof.get(dom) if prof else None}}
    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(content, args.level, args.chunk_mib << 20)
        log(f"cpu baseline in {time.perf_counter() - t:.1f} s: {cpu}")
    line = {
        "metric": "decoded GB/s, 4 GiB single .gz on 1 MI355X; batched GB/s + p99 latency",
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": f"single-member gzip, {U} B decoded: log text with 20% uniform random runs of "
                               f"64 KiB-4 MiB, zlib level {args.level}, 32 KiB window, "
                               + ("one zlib stream" if args.chunk_mib == 0 else
                                  f"{args.chunk_mib} MiB parallel chunks ended by sync flushes"),
                   "compressed_bytes": C, "decoded_bytes": U, "parallelism": "1 GPU",
                   "device_resident": True, "crc_and_isize_checked_every_step": True, "bit_exact": exact,
                   "segments": st.segments, "rounds": st.rounds, "jobs": st.jobs,
                   "schedule": tuning or "defaults"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "host_path_GBs": round(U / host_s / 1e9, 3),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
