"""Python bindings of libgzhost.so (include/gzhost.h): whole .gz files decoded member by member on
the GPU, every member's CRC-32 and size checked against its trailer. The error classes are the
reference host's gz::Fail (src/gz.rs:112-119): Corrupt (a damaged file), Retry (a GPU failure or
GPU output that fails the check: decode on the CPU), Stopped (the sink said stop)."""
import ctypes
import os

from . import gpuinflate as gi

GZ_OK, GZ_CORRUPT, GZ_RETRY, GZ_STOPPED = range(4)
LIB_NAME = "libgzhost.so"


class Report(ctypes.Structure):
    _fields_ = [("members", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("packed", ctypes.c_uint64),
                ("out", ctypes.c_uint64), ("stats", gi.Stats), ("info", ctypes.c_char * 256),
                ("note", ctypes.c_char * 256), ("error", ctypes.c_char * 256)]


class Corrupt(Exception):
    pass


class Retry(Exception):
    pass


class Stopped(Exception):
    pass


_lib = None


def load_library():
    global _lib
    if _lib is None:
        gi.load_library()  # libgzhost.so links against it
        path = os.path.join(os.path.dirname(gi.library_path()), LIB_NAME)
        lib = ctypes.CDLL(path)
        u64, vp = ctypes.c_uint64, ctypes.c_void_p
        lib.gz_header.restype = ctypes.c_int
        lib.gz_header.argtypes = [ctypes.c_char_p, u64, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.c_char_p,
                                  ctypes.c_size_t]
        lib.gz_decode.restype = ctypes.c_int
        lib.gz_decode.argtypes = [vp, ctypes.c_char_p, u64, gi.SINK, vp, ctypes.POINTER(Report)]
        _lib = lib
    return _lib


def header(data):
    """(header length, BGZF member size or 0) of the member at the start of data (src/gz.rs:62-110);
    Corrupt when it is not a valid member header. No GPU needed."""
    lib = load_library()
    n, m = ctypes.c_uint64(), ctypes.c_uint64()
    err = ctypes.create_string_buffer(256)
    data = bytes(data)
    if lib.gz_header(data, len(data), ctypes.byref(n), ctypes.byref(m), err, len(err)) != GZ_OK:
        raise Corrupt(err.value.decode(errors="replace"))
    return n.value, m.value


def decode(gpu, data, sink=None):
    """Decodes every member of the gzip file `data` on gpu (a gpuinflate.Gpu). Returns (output bytes,
    or None when a sink is given, Report). sink(bytes) -> bool: False stops (Stopped)."""
    lib = load_library()
    data = bytes(data)
    chunks, failure = [], []

    def tramp(user, ptr, n):
        try:
            b = ctypes.string_at(ptr, n)
            if sink is None:
                chunks.append(b)
                return 0
            return 0 if sink(b) else 1
        except BaseException as e:  # never unwind through the C frame
            failure.append(e)
            return 1

    cb = gi.SINK(tramp)
    rep = Report()
    with gpu._lock:
        rc = lib.gz_decode(gpu._ctx, data, len(data), cb, None, ctypes.byref(rep))
    if failure:
        raise failure[0]
    msg = rep.error.decode(errors="replace")
    if rc == GZ_CORRUPT:
        raise Corrupt(msg)
    if rc == GZ_RETRY:
        raise Retry(msg)
    if rc == GZ_STOPPED:
        raise Stopped(msg)
    return (b"".join(chunks) if sink is None else None), rep
