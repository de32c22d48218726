"""Python bindings of libgpuinflate.so, the MI355X decoder (include/gpuinflate.h).

The interface mirrors the reference host's loader (src/inflate.rs:1-168): `Gpu.load()` finds
the library (GPUZIP_INFLATE_DIR, then this package's lib/ folder), checks the ABI version for
equality (inflate.rs:124), makes one context, and `Gpu.inflate(data, hint, sink)` maps the
return codes to exceptions as inflate.rs:160-166 maps them to `Error`. There is no CPU
fallback here: without the library or a GPU, loading raises.
"""
import ctypes
import os
import threading

ABI_VERSION = 1
GI_OK, GI_ERR_DATA, GI_ERR_TRUNCATED, GI_ERR_ABORTED, GI_ERR_GPU, GI_ERR_ARG = range(6)
CODE_NAMES = {0: "ok", 1: "data", 2: "truncated", 3: "aborted", 4: "gpu", 5: "arg"}

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_NAME = "libgpuinflate.so"

# every symbol include/gpuinflate.h declares
EXPORTS = ["gi_abi_version", "gi_create", "gi_inflate", "gi_last_error", "gi_device_name", "gi_destroy",
           "gi_last_crc32", "gi_inflate_device", "gi_kernel_times", "gi_batch_create", "gi_batch_run",
           "gi_batch_submit", "gi_batch_wait", "gi_batch_workers", "gi_batch_kernel_times", "gi_batch_destroy",
           "gi_rescue_headers",
           "gi_rescue_headers_noalloc", "gi_inflate_many", "gi_serve_create", "gi_serve_submit", "gi_serve_wait",
           "gi_serve_in_use", "gi_serve_destroy"]


class Config(ctypes.Structure):
    _fields_ = [("size", ctypes.c_uint32), ("device", ctypes.c_int32), ("segment_out_mb", ctypes.c_uint32),
                ("segment_in_kb", ctypes.c_uint32)]


class ConfigTuned(ctypes.Structure):
    """gi_config_tuned: a gi_config with the decoder's schedule after it (0 keeps a default)"""
    _fields_ = [("base", Config), ("chunk_kib", ctypes.c_uint32), ("probe_spacing_kbit", ctypes.c_uint32),
                ("probe_lead_bits", ctypes.c_uint32), ("segment_in_max_mb", ctypes.c_uint32),
                ("lanes_per_group", ctypes.c_uint32), ("persistent_per_cu", ctypes.c_uint32)]


# the schedule's fields, and their ranges (include/gpuinflate.h)
TUNING_RANGES = {"chunk_kib": (4, 64), "probe_spacing_kbit": (64, 1024), "probe_lead_bits": (1024, 8192),
                 "segment_in_max_mb": (64, 1024), "lanes_per_group": (16, 64), "persistent_per_cu": (1, 16)}
LANES_PER_GROUP = (16, 32, 64)  # the values lanes_per_group takes within its range


def make_config(device=0, segment_out_mb=0, segment_in_kb=0, tuning=None):
    """a gi_config, or a gi_config_tuned when `tuning` (a dict of TUNING_RANGES' fields) is given"""
    if not tuning:
        return Config(ctypes.sizeof(Config), device, segment_out_mb, segment_in_kb)
    unknown = set(tuning) - set(TUNING_RANGES)
    if unknown:
        raise ValueError(f"unknown schedule fields {sorted(unknown)}")
    cfg = ConfigTuned(Config(ctypes.sizeof(ConfigTuned), device, segment_out_mb, segment_in_kb))
    for k, v in tuning.items():
        setattr(cfg, k, int(v))
    return cfg


class Stats(ctypes.Structure):
    _fields_ = [("in_used", ctypes.c_uint64), ("out_size", ctypes.c_uint64), ("segments", ctypes.c_uint32),
                ("rounds", ctypes.c_uint32), ("jobs", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("upload_s", ctypes.c_double), ("find_s", ctypes.c_double), ("decode_s", ctypes.c_double),
                ("resolve_s", ctypes.c_double), ("download_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("launches", ctypes.c_uint64), ("seconds", ctypes.c_double)]


class Job(ctypes.Structure):
    _fields_ = [("d_in", ctypes.c_void_p), ("in_size", ctypes.c_uint64), ("d_out", ctypes.c_void_p),
                ("out_cap", ctypes.c_uint64), ("out_hint", ctypes.c_uint64), ("user", ctypes.c_void_p),
                ("out_size", ctypes.c_uint64), ("in_used", ctypes.c_uint64), ("crc32", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("queued_s", ctypes.c_double), ("seconds", ctypes.c_double)]


class Stream(ctypes.Structure):  # gi_stream
    _fields_ = [("offset", ctypes.c_uint64), ("size", ctypes.c_uint64), ("out_offset", ctypes.c_uint64),
                ("out_room", ctypes.c_uint64), ("out_size", ctypes.c_uint64), ("in_used", ctypes.c_uint64),
                ("crc32", ctypes.c_uint32), ("status", ctypes.c_int32)]


class Request(ctypes.Structure):  # gi_request
    _fields_ = [("in_", ctypes.c_void_p), ("in_size", ctypes.c_uint64), ("out", ctypes.c_void_p),
                ("out_cap", ctypes.c_uint64), ("gzip", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("user", ctypes.c_void_p), ("out_size", ctypes.c_uint64), ("crc32", ctypes.c_uint32),
                ("status", ctypes.c_int32), ("admitted_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("error", ctypes.c_char * 96)]


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)


class Error(Exception):
    def __init__(self, code, message):
        super().__init__(f"{CODE_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


class DataError(Error):
    pass


class TruncatedError(Error):
    pass


class AbortedError(Error):
    pass


class GpuError(Error):
    pass


class ArgError(Error):
    pass


_ERRORS = {GI_ERR_DATA: DataError, GI_ERR_TRUNCATED: TruncatedError, GI_ERR_ABORTED: AbortedError,
           GI_ERR_GPU: GpuError, GI_ERR_ARG: ArgError}


def library_path():
    """The library's path: $GPUZIP_INFLATE_DIR first, then this package's lib/ (dylib.rs:56-77)."""
    cands = []
    d = os.environ.get("GPUZIP_INFLATE_DIR")
    if d:
        cands.append(os.path.join(d, LIB_NAME))
    cands.append(os.path.join(LIB_DIR, LIB_NAME))
    for c in cands:
        if os.path.exists(c):
            return c
    raise FileNotFoundError(f"{LIB_NAME} not found in {cands}: build it with `python -m gpuzip_amd.build`")


_lib = None
_lib_lock = threading.Lock()


def load_library():
    """Loads libgpuinflate.so and declares its signatures (no GPU needed)."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        lib = ctypes.CDLL(library_path())
        u64, u32, vp = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
        lib.gi_abi_version.restype = u32
        lib.gi_abi_version.argtypes = []
        lib.gi_create.restype = ctypes.c_int
        lib.gi_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]
        lib.gi_inflate.restype = ctypes.c_int
        lib.gi_inflate.argtypes = [vp, ctypes.c_char_p, u64, u64, SINK, vp, ctypes.POINTER(Stats)]
        lib.gi_last_error.restype = ctypes.c_char_p
        lib.gi_last_error.argtypes = [vp]
        lib.gi_device_name.restype = ctypes.c_char_p
        lib.gi_device_name.argtypes = [vp]
        lib.gi_destroy.restype = None
        lib.gi_destroy.argtypes = [vp]
        lib.gi_last_crc32.restype = u32
        lib.gi_last_crc32.argtypes = [vp]
        lib.gi_inflate_device.restype = ctypes.c_int
        lib.gi_inflate_device.argtypes = [vp, vp, u64, vp, u64, u64, ctypes.POINTER(u64), ctypes.POINTER(u32),
                                          ctypes.POINTER(Stats)]
        lib.gi_kernel_times.restype = ctypes.c_int
        lib.gi_kernel_times.argtypes = [vp, ctypes.POINTER(KernelTime), ctypes.c_int]
        lib.gi_rescue_headers.restype = ctypes.c_int
        lib.gi_rescue_headers.argtypes = [vp, ctypes.c_char_p, u64, u64, vp, u64, ctypes.POINTER(u64)]
        lib.gi_rescue_headers_noalloc.restype = ctypes.c_int
        lib.gi_rescue_headers_noalloc.argtypes = [ctypes.c_int, vp, u64, u64, vp, u64, ctypes.POINTER(u64), ctypes.c_char_p,
                                                  ctypes.c_size_t]
        lib.gi_batch_create.restype = ctypes.c_int
        lib.gi_batch_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(vp), ctypes.c_char_p,
                                        ctypes.c_size_t]
        lib.gi_batch_run.restype = ctypes.c_int
        lib.gi_batch_run.argtypes = [vp, ctypes.POINTER(Job), u64]
        lib.gi_batch_submit.restype = ctypes.c_int
        lib.gi_batch_submit.argtypes = [vp, ctypes.POINTER(Job)]
        lib.gi_batch_wait.restype = ctypes.POINTER(Job)
        lib.gi_batch_wait.argtypes = [vp, ctypes.c_double]
        lib.gi_batch_workers.restype = ctypes.c_int
        lib.gi_batch_workers.argtypes = [vp]
        lib.gi_batch_kernel_times.restype = ctypes.c_int
        lib.gi_batch_kernel_times.argtypes = [vp, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(u64),
                                              ctypes.POINTER(u64), ctypes.c_int]
        lib.gi_inflate_many.restype = ctypes.c_int
        lib.gi_inflate_many.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(Stream), u64, vp, u64, ctypes.POINTER(Stats)]
        lib.gi_serve_create.restype = ctypes.c_int
        lib.gi_serve_create.argtypes = [ctypes.POINTER(Config), ctypes.c_int, u64, ctypes.POINTER(vp), ctypes.c_char_p,
                                        ctypes.c_size_t]
        lib.gi_serve_submit.restype = ctypes.c_int
        lib.gi_serve_submit.argtypes = [vp, ctypes.POINTER(Request)]
        lib.gi_serve_wait.restype = ctypes.POINTER(Request)
        lib.gi_serve_wait.argtypes = [vp, ctypes.c_double]
        lib.gi_serve_in_use.restype = u64
        lib.gi_serve_in_use.argtypes = [vp, ctypes.POINTER(u64)]
        lib.gi_serve_destroy.restype = None
        lib.gi_serve_destroy.argtypes = [vp]
        lib.gi_batch_destroy.restype = None
        lib.gi_batch_destroy.argtypes = [vp]
        _lib = lib
        return lib


def _raise(code, msg):
    raise _ERRORS.get(code, Error)(code, msg)


class Gpu:
    """One decoder context on one GPU (inflate.rs:41-168). Not for use from two threads at once;
    calls are serialised with a lock as the reference host does (inflate.rs:60, 152)."""

    def __init__(self, device=0, segment_out_mb=0, segment_in_kb=0, tuning=None):
        lib = load_library()
        if lib.gi_abi_version() != ABI_VERSION:
            raise GpuError(GI_ERR_GPU, f"{LIB_NAME} has ABI {lib.gi_abi_version()}, expected {ABI_VERSION}")
        self._lib = lib
        self._lock = threading.Lock()
        cfg = make_config(device, segment_out_mb, segment_in_kb, tuning)
        ctx = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, len(err))
        if rc != GI_OK:
            _raise(rc, err.value.decode(errors="replace"))
        self._ctx = ctx
        self.device = lib.gi_device_name(ctx).decode(errors="replace")

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.gi_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_error(self):
        return self._lib.gi_last_error(self._ctx).decode(errors="replace")

    def last_crc32(self):
        return self._lib.gi_last_crc32(self._ctx)

    def rescue_headers(self, data, start):
        """The rescue's header search (gpuinflate/rescue.cpp:597-626) on the GPU: the confirmed
        dynamic-block headers (bit positions) in [start, 8 * len(data) - 128), in order."""
        import numpy as np
        data = bytes(data)
        cap = max(1024, len(data) // 64)
        out = np.zeros(cap, np.uint64)
        n = ctypes.c_uint64()
        with self._lock:
            rc = self._lib.gi_rescue_headers(self._ctx, data, len(data), start, out.ctypes.data, cap, ctypes.byref(n))
        if rc != GI_OK:
            _raise(rc, self.last_error())
        if n.value > cap:
            raise Error(GI_ERR_ARG, "more headers than the output holds")
        return out[: n.value].tolist()

    def kernel_times(self):
        """{kernel name: (launches, seconds)} of the last call, timed with HIP events"""
        arr = (KernelTime * 32)()
        n = self._lib.gi_kernel_times(self._ctx, arr, 32)
        return {arr[i].name.decode(): (arr[i].launches, arr[i].seconds) for i in range(n)}

    def inflate(self, data, hint=0, sink=None):
        """Decodes the raw Deflate stream at the start of `data`. Returns (output bytes or None
        when a sink is given, Stats). sink(bytes) -> bool: False stops the decode (aborted)."""
        data = bytes(data)
        chunks = []
        failure = []

        def tramp(user, ptr, n):
            try:
                b = ctypes.string_at(ptr, n)
                if sink is None:
                    chunks.append(b)
                    return 0
                return 0 if sink(b) else 1
            except BaseException as e:  # never unwind through the C frame (inflate.rs:99-107)
                failure.append(e)
                return 1

        cb = SINK(tramp)
        st = Stats()
        with self._lock:
            rc = self._lib.gi_inflate(self._ctx, data, len(data), hint, cb, None, ctypes.byref(st))
        if failure:
            raise failure[0]
        if rc != GI_OK:
            _raise(rc, self.last_error())
        return (b"".join(chunks) if sink is None else None), st

    def inflate_device(self, d_in, in_size, d_out, out_cap, hint=0):
        """Device-resident decode: d_in/d_out are device addresses (ints). Returns
        (out_size, crc32, Stats)."""
        st = Stats()
        osz = ctypes.c_uint64()
        crc = ctypes.c_uint32()
        with self._lock:
            rc = self._lib.gi_inflate_device(self._ctx, ctypes.c_void_p(d_in), in_size, ctypes.c_void_p(d_out), out_cap,
                                             hint, ctypes.byref(osz), ctypes.byref(crc), ctypes.byref(st))
        if rc != GI_OK:
            _raise(rc, self.last_error())
        return osz.value, crc.value, st


    def inflate_many(self, data, spans):
        """gi_inflate_many: the raw Deflate streams data[offset:offset + size] for (offset, size,
        out_room) in spans, decoded on this context as batch segments. Returns (the streams'
        results as gi_stream structures, their outputs as bytes or None where a stream failed,
        the call's return code)."""
        data = bytes(data)
        arr = (Stream * max(1, len(spans)))()
        at = 0
        for i, (off, size, room) in enumerate(spans):
            arr[i].offset, arr[i].size, arr[i].out_offset, arr[i].out_room = off, size, at, room
            at += room
        out = ctypes.create_string_buffer(max(1, at))
        st = Stats()
        with self._lock:
            rc = self._lib.gi_inflate_many(self._ctx, data, arr, len(spans), out, at, ctypes.byref(st))
        if rc == GI_ERR_ARG:
            _raise(rc, self.last_error())
        res = [arr[i] for i in range(len(spans))]
        outs = [out.raw[r.out_offset:r.out_offset + r.out_size] if r.status == GI_OK else None for r in res]
        return res, outs, rc


class Batch:
    """A pool of `workers` decoder contexts on one GPU (include/gpuinflate.h, gi_batch_*): run() a
    whole batch of device-resident streams, or submit()/wait() one stream at a time."""

    def __init__(self, workers=2, device=0, segment_out_mb=0, segment_in_kb=0, tuning=None):
        lib = load_library()
        self._lib = lib
        cfg = make_config(device, segment_out_mb, segment_in_kb, tuning)
        b = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.gi_batch_create(ctypes.byref(cfg), workers, ctypes.byref(b), err, len(err))
        if rc != GI_OK:
            _raise(rc, err.value.decode(errors="replace"))
        self._b = b
        self.workers = lib.gi_batch_workers(b)

    def run(self, jobs):
        """jobs: a ctypes array of Job; returns the first failed job's status, or GI_OK"""
        return self._lib.gi_batch_run(self._b, jobs, len(jobs))

    def submit(self, job):
        rc = self._lib.gi_batch_submit(self._b, ctypes.byref(job))
        if rc != GI_OK:
            _raise(rc, "submit")

    def wait(self, timeout=1.0):
        """the address of the next finished submitted Job, or None"""
        p = self._lib.gi_batch_wait(self._b, timeout)
        return ctypes.addressof(p.contents) if p else None

    def kernel_times(self, reset=False):
        """({kernel name: (launches, seconds)}, batch segments, streams decoded alone) since the
        last reset, summed over the workers"""
        arr = (KernelTime * 32)()
        seg, alone = ctypes.c_uint64(), ctypes.c_uint64()
        n = self._lib.gi_batch_kernel_times(self._b, arr, 32, ctypes.byref(seg), ctypes.byref(alone), 1 if reset else 0)
        return {arr[i].name.decode(): (arr[i].launches, arr[i].seconds) for i in range(n)}, seg.value, alone.value

    def close(self):
        if getattr(self, "_b", None):
            self._lib.gi_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """A serving loop on one GPU (gi_serve_*): submit() gzip members or raw streams from host memory
    as they arrive; wait() hands back finished ones. Admission by the HBM budget, decoding by
    `workers` batch workers."""

    def __init__(self, budget, workers=2, device=0, segment_out_mb=0, segment_in_kb=0, tuning=None):
        lib = load_library()
        self._lib = lib
        cfg = make_config(device, segment_out_mb, segment_in_kb, tuning)
        p = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = lib.gi_serve_create(ctypes.cast(ctypes.byref(cfg), ctypes.POINTER(Config)), workers, budget,
                                 ctypes.byref(p), err, len(err))
        if rc != GI_OK:
            _raise(rc, err.value.decode(errors="replace"))
        self._s = p
        self._live = {}  # request address -> (request, input, output): kept alive until handed back

    @staticmethod
    def prepare(data, out_cap, gzip=True, tag=None):
        """a request and its buffers, ready for submit() (preparing copies the input and makes the
        output buffer, so that submitting costs no more than the call)"""
        data = bytes(data)
        out = ctypes.create_string_buffer(max(1, out_cap))
        r = Request()
        r.in_ = ctypes.cast(ctypes.c_char_p(data), ctypes.c_void_p)
        r.in_size = len(data)
        r.out = ctypes.cast(out, ctypes.c_void_p)
        r.out_cap = out_cap
        r.gzip = 1 if gzip else 0
        return (r, data, out, tag)

    def submit(self, data, out_cap=0, gzip=True, tag=None):
        """queues one request (data: bytes, or what prepare() made); returns its gi_request, whose
        results are filled in when wait() hands it back"""
        prepared = data if isinstance(data, tuple) else self.prepare(data, out_cap, gzip, tag)
        r = prepared[0]
        self._live[ctypes.addressof(r)] = prepared
        rc = self._lib.gi_serve_submit(self._s, ctypes.byref(r))
        if rc != GI_OK:
            del self._live[ctypes.addressof(r)]
            _raise(rc, "submit")
        return r

    def wait(self, timeout=1.0):
        """(request, output bytes or None, tag) of the next finished request, or None"""
        p = self._lib.gi_serve_wait(self._s, timeout)
        if not p:
            return None
        r, data, out, tag = self._live.pop(ctypes.addressof(p.contents))
        return r, (out.raw[: r.out_size] if r.status == GI_OK else None), tag

    def in_use(self):
        """(device bytes held by admitted requests, the most ever held)"""
        peak = ctypes.c_uint64()
        now = self._lib.gi_serve_in_use(self._s, ctypes.byref(peak))
        return now, peak.value

    def close(self):
        if getattr(self, "_s", None):
            self._lib.gi_serve_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def rescue_headers_noalloc(d_in, size, start, device=0, cap=None):
    """gi_rescue_headers_noalloc: the rescue's header search over size bytes at device address
    d_in, the library allocating nothing; the confirmed headers (bit positions) in order"""
    import numpy as np
    lib = load_library()
    cap = cap if cap is not None else max(1024, size // 64)
    out = np.zeros(cap, np.uint64)
    n = ctypes.c_uint64()
    err = ctypes.create_string_buffer(256)
    rc = lib.gi_rescue_headers_noalloc(device, ctypes.c_void_p(d_in), size, start, out.ctypes.data, cap, ctypes.byref(n),
                                       err, len(err))
    if rc != GI_OK:
        _raise(rc, err.value.decode(errors="replace"))
    if n.value > cap:
        raise Error(GI_ERR_ARG, "more headers than the output holds")
    return out[: n.value].tolist()

