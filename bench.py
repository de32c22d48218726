"""Benchmark: decoded GB/s of one 4 GiB single-member gzip stream on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size-mib 4096]

Workload: 4 GiB (4,294,967,296 B) of synthetic log text interleaved with seeded uniform random
runs of 64 KiB-4 MiB (20% of the bytes), one gzip member, zlib level 6 with a 32 KiB window,
compressed in parallel 64 MiB chunks chained by their 32 KiB dictionaries (one valid stream;
ISIZE wraps to 0). A step is one decode of the whole member: the compressed body is resident in
device memory before timing starts, the output goes to device memory, and the GPU's CRC-32 and
the size are checked against the member's trailer every step. After the timed steps the whole
output is compared byte for byte with the generated original.

Prints one JSON line. `roofline` is the dominant kernel's (the one with the most GPU time in the
timed steps, from HIP events recorded around every launch on the library's stream); its
algorithmic bytes are the stream's C + U per step (compressed bytes read once, decoded bytes
written once). `cpu_baseline` times the reference's own sequential decoder (gpuinflate/
inflate.cuh built from its sources into oracle/_ref) on one host core over a 1 GiB sample of the
same content.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(content, level, chunk):
    """the reference decoder on one core over a bounded sample of the same content"""
    import ctypes

    import numpy as np

    from gpuzip_amd import datagen
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib

    n = min(content.size, 1 << 30)
    sample = content[:n]
    s = datagen.deflate(sample, level, 0, chunk=chunk)
    lib = oracle_lib.ref()
    kind, fn = ("reference", lib.ref_inflate) if lib is not None else ("port", oracle_lib.oracle().or_inflate)
    out = np.empty(n, np.uint8)
    ol, iu, tb = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int()
    sb = s.tobytes()
    t = time.perf_counter()
    rc = fn(sb, len(sb), out.ctypes.data, n, ctypes.byref(ol), ctypes.byref(iu), ctypes.byref(tb))
    dt = time.perf_counter() - t
    ok = rc == 0 and ol.value == n and np.array_equal(out, sample)
    return {"value": round(n / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"first {n >> 20} MiB of the same content, zlib level {level}, {len(sb)} B compressed, "
                      f"decoded sequentially in {dt:.2f} s, output {'bit-exact' if ok else 'WRONG'}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size-mib", type=int, default=4096)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--chunk-mib", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="one device-to-device copy of this many MiB before the steps: the known-size "
                         "kernel that scales rocprofv3's FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): this decoder has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from gpuzip_amd import datagen
    from gpuzip_amd.gpuinflate import Gpu
    from gpuzip_amd.hipmem import D2D, DeviceBuffer, check, hip

    U = args.size_mib << 20
    t = time.perf_counter()
    content = datagen.mixed(U)
    log(f"generated {U} B in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    body = datagen.deflate(content, args.level, 0, chunk=args.chunk_mib << 20)
    crc = datagen.crc32(content)
    log(f"compressed to {body.size} B (C/U {body.size / U:.3f}) in {time.perf_counter() - t:.1f} s")
    # the member: header (10 B), body, trailer (CRC-32, ISIZE = size mod 2^32)
    isize = U & 0xFFFFFFFF
    C = body.size

    g = Gpu(device=0)
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
    for k in range(args.steps):
        st = step()
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
    for name, (launches, sec) in sorted(ktimes.items(), key=lambda kv: -kv[1][1]):
        log(f"  {name:12s} {launches:6d} launches {sec * 1e3 / args.steps:9.2f} ms/step")

    out = d_out.to_host()
    exact = bool(np.array_equal(out, content))
    del out
    if not exact:
        raise SystemExit("output differs from the original")
    log("output bit-exact with the original")

    dom, (dl, ds) = max(ktimes.items(), key=lambda kv: kv[1][1])
    alg = (C + U) * args.steps
    achieved = alg / ds / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": dom,
                "launches_per_step": dl / args.steps, "kernel_ms_per_step": round(ds / args.steps * 1e3, 3),
                "alg_bytes_per_step": C + U}
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
                               f"{args.chunk_mib} MiB parallel chunks",
                   "compressed_bytes": C, "decoded_bytes": U, "parallelism": "1 GPU",
                   "device_resident": True, "crc_and_isize_checked_every_step": True, "bit_exact": exact,
                   "segments": st.segments, "rounds": st.rounds, "jobs": st.jobs},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
