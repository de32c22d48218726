"""The rescue of damaged gzip files and raw Deflate streams on the GPU (libgzrescue.so, behind
include/gzrescue.h), with the host interface of the reference's loader (src/rescue.rs:131-262):
`get()` loads the library once and checks its ABI version; `Rescue.rescue(input, fill, sink)`
and `Rescue.rescue_raw(input, expected, fill, crs, cr_budget, sink)` pass the output in order to
`sink(data, mask)` (return False to stop) and return a Report of statistics, pieces and notes, or
raise RescueError with the reference's messages. There is no CPU fallback: a missing library or
GPU raises.

    from gpuzip_amd import rescue
    rep = rescue.get().rescue(open("damaged.gz", "rb").read(), 0, lambda d, m: True)
"""
import ctypes
import os
import threading
from dataclasses import dataclass, field

ABI_VERSION = 4
LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_NAME = "libgzrescue.so"

# mask values (rescue.h:40-43)
UNKNOWN, EXACT, SUSPECT, INFERRED = 0, 1, 2, 3
GR_OK, GR_ERR_NO_DEFLATE, GR_ERR_ABORTED, GR_ERR_ARG, GR_ERR_MEMORY, GR_ERR_TOO_BIG, GR_ERR_INTERNAL = range(7)

u8, u32, u64, i32, i64 = ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64


class Options(ctypes.Structure):  # gr_options
    _fields_ = [("size", u32), ("fill", u8), ("no_infer", u8), ("cr_lost", u8), ("crs_expected", u8),
                ("max_output", u64), ("cr_budget", u64)]


class PieceRaw(ctypes.Structure):  # gr_piece
    _fields_ = [("out_offset", u64), ("length", u64), ("decoded", u64), ("known", u64), ("in_byte_start", u64),
                ("in_byte_end", u64), ("suspect", i32), ("status", i32), ("blocks", u32), ("symbols", u32),
                ("why", ctypes.c_char_p)]


class Stats(ctypes.Structure):  # gr_stats
    _fields_ = [("input_size", u64), ("output_size", u64), ("exact_bytes", u64), ("suspect_bytes", u64),
                ("unknown_bytes", u64), ("headers", u32), ("pieces", u32), ("damaged", u32), ("members", u32),
                ("crc_check", i32), ("length_off", i64), ("truncated", i32), ("header_damaged", i32),
                ("input_end", u64), ("inferred_bytes", u64), ("ascii_undone", i32), ("cr_restored", u32),
                ("cr_decoded", u64), ("hidden_damage", i32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KernelTime(ctypes.Structure):  # gzr_kernel_time
    _fields_ = [("name", ctypes.c_char_p), ("launches", u64), ("seconds", ctypes.c_double)]


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(u8), ctypes.POINTER(u8), ctypes.c_size_t)


@dataclass
class Piece:
    out_offset: int
    length: int
    decoded: int
    known: int
    in_byte_start: int
    in_byte_end: int
    suspect: bool
    why: str
    blocks: int
    symbols: int


@dataclass
class Report:
    stats: dict
    pieces: list
    notes: list
    phases: dict = field(default_factory=dict)   # seconds by step of the rescue
    kernels: dict = field(default_factory=dict)  # {kernel: (launches, GPU seconds)}
    stored_bytes: int = 0                         # output bytes copied from stored blocks by their kernel


class RescueError(RuntimeError):
    pass


# the reference's output limit when none is set: 100 times the input, at least 4 GiB (rescue.rs:273-278)
def default_max_output(input_len):
    return max(input_len * 100, 4 << 30)


class Rescue:
    def __init__(self, path=None):
        self.path = path or os.path.join(LIB_DIR, LIB_NAME)
        if not os.path.exists(self.path):
            raise RescueError(f"{LIB_NAME} not found at {self.path}: build it with python -m gpuzip_amd.build")
        lib = ctypes.CDLL(self.path)
        lib.gr_abi_version.restype = u32
        if lib.gr_abi_version() != ABI_VERSION:
            raise RescueError(f"{LIB_NAME} speaks version {lib.gr_abi_version()} of the interface, this host {ABI_VERSION}")
        lib.gr_rescue.argtypes = [ctypes.c_void_p, u64, ctypes.POINTER(Options), SINK, ctypes.c_void_p,
                                  ctypes.POINTER(ctypes.c_void_p)]
        lib.gr_rescue_raw.argtypes = [ctypes.c_void_p, u64, u64, ctypes.c_int, u32, ctypes.POINTER(Options), SINK,
                                      ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        lib.gr_result_stats.restype = ctypes.POINTER(Stats)
        lib.gr_result_stats.argtypes = [ctypes.c_void_p]
        lib.gr_result_pieces.restype = u32
        lib.gr_result_pieces.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(PieceRaw))]
        lib.gr_result_notes.restype = u32
        lib.gr_result_notes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p))]
        lib.gr_free.argtypes = [ctypes.c_void_p]
        lib.gzr_result_phases.restype = u32
        lib.gzr_result_phases.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.POINTER(ctypes.c_double), u32]
        lib.gzr_result_kernels.restype = ctypes.c_int
        lib.gzr_result_kernels.argtypes = [ctypes.c_void_p, ctypes.POINTER(KernelTime), ctypes.c_int,
                                           ctypes.POINTER(u64)]
        lib.gzr_last_error.restype = ctypes.c_char_p
        lib.gzr_set_device.argtypes = [ctypes.c_int]
        self._lib = lib

    def set_device(self, device):
        if self._lib.gzr_set_device(device) != GR_OK:
            raise RescueError(f"no device {device}")

    def rescue(self, data, fill=0, sink=None, max_output=None, no_infer=False, cr_lost=False):
        """the gzip file in `data`"""
        return self._run(data, None, fill, 0, 0, sink, max_output, no_infer, cr_lost)

    def rescue_raw(self, data, expected=(None, None), fill=0, crs=0, cr_budget=0, sink=None, max_output=None,
                   no_infer=False, cr_lost=False):
        """a raw Deflate stream at data[0]; expected: (size or None, CRC-32 or None) in place of a trailer"""
        return self._run(data, expected, fill, crs, cr_budget, sink, max_output, no_infer, cr_lost)

    def _run(self, data, raw, fill, crs, cr_budget, sink, max_output, no_infer, cr_lost):
        lib = self._lib
        buf = data if isinstance(data, (bytes, bytearray)) else bytes(data)
        n = len(buf)
        limit = default_max_output(n) if max_output is None else max_output
        opt = Options(ctypes.sizeof(Options), fill, 1 if no_infer else 0, 1 if cr_lost else 0, crs, limit, cr_budget)
        failed = []

        def tramp(user, d, m, k):
            if sink is None:
                return 0
            try:
                return 0 if sink(ctypes.string_at(d, k), ctypes.string_at(m, k)) else 1
            except Exception as e:  # a sink that raises stops the rescue, as a panic does in the reference host
                failed.append(e)
                return 1

        cb = SINK(tramp)
        res = ctypes.c_void_p()
        src = (ctypes.c_char * n).from_buffer(buf) if isinstance(buf, bytearray) else ctypes.c_char_p(buf)
        if raw is None:
            rc = lib.gr_rescue(ctypes.cast(src, ctypes.c_void_p), n, ctypes.byref(opt), cb, None, ctypes.byref(res))
        else:
            size, crc = raw
            rc = lib.gr_rescue_raw(ctypes.cast(src, ctypes.c_void_p), n, (1 << 64) - 1 if size is None else size,
                                   0 if crc is None else 1, crc or 0, ctypes.byref(opt), cb, None, ctypes.byref(res))
        if not res.value:
            raise RescueError({
                GR_ERR_NO_DEFLATE: "no Deflate data could be recognised",
                GR_ERR_MEMORY: "out of memory",
                GR_ERR_TOO_BIG: f"the output would be larger than {limit} bytes ({limit / max(n, 1):.0f} times the "
                                f"input: a decompression bomb, or data that compresses extremely well)",
                GR_ERR_INTERNAL: f"gzrescue failed: {lib.gzr_last_error().decode(errors='replace')}",
            }.get(rc, f"gzrescue failed ({rc})"))
        try:
            stats = lib.gr_result_stats(res).contents.as_dict()
            pp = ctypes.POINTER(PieceRaw)()
            pieces = []
            for k in range(lib.gr_result_pieces(res, ctypes.byref(pp))):
                p = pp[k]
                pieces.append(Piece(p.out_offset, p.length, p.decoded, p.known, p.in_byte_start, p.in_byte_end,
                                    p.suspect != 0, (p.why or b"").decode(errors="replace"), p.blocks, p.symbols))
            nn = ctypes.POINTER(ctypes.c_char_p)()
            notes = [nn[k].decode(errors="replace") for k in range(lib.gr_result_notes(res, ctypes.byref(nn)))]
            names, secs = (ctypes.c_char_p * 64)(), (ctypes.c_double * 64)()
            phases = {names[i].decode(): secs[i] for i in range(min(64, lib.gzr_result_phases(res, names, secs, 64)))}
            kt, stored = (KernelTime * 32)(), u64()
            kernels = {kt[i].name.decode(): (kt[i].launches, kt[i].seconds)
                       for i in range(min(32, lib.gzr_result_kernels(res, kt, 32, ctypes.byref(stored))))}
        finally:
            lib.gr_free(res)
        if rc == GR_ERR_ABORTED:
            if failed:
                raise failed[0]
            raise RescueError("stopped while writing the output")
        return Report(stats, pieces, notes, phases, kernels, stored.value)


_lock = threading.Lock()
_lib = None


def get():
    """the library, loaded once per process on first use (rescue.rs:130-134)"""
    global _lib
    with _lock:
        if _lib is None:
            _lib = Rescue()
        return _lib
