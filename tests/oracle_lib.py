"""ctypes access to the test oracle (oracle/liboracle.so), to the reference's own decoder built
from its sources (oracle/_ref/libgiref.so, when present) and to the host build of the decoding
core (gpuzip_amd/lib/libgzhostcheck.so). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
u64, u32, c_int = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int


class OrStart(ctypes.Structure):
    _fields_ = [("bit", u64), ("header", u64), ("mode", ctypes.c_uint8), ("final", ctypes.c_int8)]


class OrResult(ctypes.Structure):
    _fields_ = [("end_bit", u64), ("eob_bit", u64), ("resume", OrStart), ("eob_out", u32), ("out_len", u32),
                ("status", ctypes.c_int32), ("landed", ctypes.c_int32), ("end_final", ctypes.c_int32),
                ("blocks", u32), ("last_block_bit", u64), ("last_block_out", u32), ("symbols", u32)]


_cache = {}


def _load(name, path):
    if name not in _cache:
        _cache[name] = ctypes.CDLL(path) if os.path.exists(path) else None
    return _cache[name]


def oracle():
    lib = _load("oracle", os.path.join(ROOT, "oracle", "liboracle.so"))
    if lib is None:
        raise FileNotFoundError("oracle/liboracle.so: run `make -C oracle`")
    lib.or_crc32.restype = u32
    lib.or_crc32.argtypes = [u32, ctypes.c_char_p, ctypes.c_size_t]
    lib.or_header_prefilter.argtypes = [u64]
    lib.or_header_prefilter128.argtypes = [u64, u64]
    lib.or_plausible_dynamic_header.argtypes = [ctypes.c_char_p, u64, u64]
    lib.or_bits64.restype = u64
    lib.or_bits64.argtypes = [ctypes.c_char_p, u64, u64]
    lib.or_inflate.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                               ctypes.POINTER(c_int)]
    lib.or_find_headers.restype = u64
    lib.or_find_headers.argtypes = [ctypes.c_char_p, u64, u64, u64, ctypes.c_void_p, u64]
    lib.or_run.restype = OrResult
    lib.or_run.argtypes = [ctypes.c_char_p, u64, OrStart, ctypes.POINTER(OrStart), c_int, ctypes.c_void_p, u32, c_int, u64]
    return lib


def ref():
    """the reference's decoder (gpuinflate/inflate.cuh) behind oracle/ref_shim.cpp, or None"""
    lib = _load("ref", os.path.join(ROOT, "oracle", "_ref", "libgiref.so"))
    if lib is None:
        return None
    lib.ref_inflate.argtypes = [ctypes.c_char_p, u64, ctypes.c_void_p, u64, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                ctypes.POINTER(c_int)]
    lib.ref_plausible_dynamic_header.argtypes = [ctypes.c_char_p, u64, u64]
    lib.ref_header_prefilter.argtypes = [u64]
    lib.ref_header_prefilter128.argtypes = [u64, u64]
    lib.ref_run.restype = None
    lib.ref_run.argtypes = [ctypes.c_char_p, u64, ctypes.POINTER(OrStart), ctypes.POINTER(OrStart), c_int,
                            ctypes.c_void_p, u32, c_int, u64, ctypes.POINTER(OrResult)]
    return lib


def hostcheck():
    # GZ_HOSTCHECK_LIB: another build of the same source (the sanitizer test's)
    path = os.environ.get("GZ_HOSTCHECK_LIB") or os.path.join(ROOT, "gpuzip_amd", "lib", "libgzhostcheck.so")
    lib = _load("hostcheck", path)
    if lib is None:
        raise FileNotFoundError("gpuzip_amd/lib/libgzhostcheck.so: run `python -m gpuzip_amd.build`")
    lib.hc_inflate.argtypes = [ctypes.c_char_p, u64, u64, u64, u64, ctypes.c_void_p, u64, ctypes.POINTER(u64),
                               ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.c_char_p,
                               ctypes.c_size_t]
    lib.hc_find.restype = u64
    lib.hc_find.argtypes = [ctypes.c_char_p, u64, u64, u64, ctypes.c_void_p, u64]
    lib.hc_prefilter.restype = u64
    lib.hc_prefilter.argtypes = [ctypes.c_char_p, u64, u64, u64, ctypes.c_void_p, u64]
    lib.hc_sizeof.restype = u32
    lib.hc_set_schedule.restype = None
    lib.hc_set_schedule.argtypes = [u32, u32, u32, u32]
    return lib


def _whole(fn, data, max_out):
    out = np.empty(max(max_out, 1), np.uint8)
    ol, iu, tb = u64(), u64(), c_int()
    rc = fn(bytes(data), len(data), out.ctypes.data, max_out, ctypes.byref(ol), ctypes.byref(iu), ctypes.byref(tb))
    return rc, bytes(out[: ol.value]), iu.value, bool(tb.value)


def oracle_inflate(data, max_out=None):
    """(status 0/1/2, bytes, in_used, too_big) -- the sequential decode of gi_test.cpp:47-91"""
    if max_out is None:
        max_out = max(64 << 20, len(data) * 16)
    return _whole(oracle().or_inflate, data, max_out)


def ref_inflate(data, max_out=None):
    lib = ref()
    if lib is None:
        return None
    if max_out is None:
        max_out = max(64 << 20, len(data) * 16)
    return _whole(lib.ref_inflate, data, max_out)


def hc_inflate(data, hint=0, seg_out=256 << 20, seg_in_kb=0, max_out=None):
    if max_out is None:
        max_out = max(64 << 20, len(data) * 16)
    lib = hostcheck()
    out = np.empty(max(max_out, 1), np.uint8)
    ol, iu, sg, rd = u64(), u64(), u32(), u32()
    err = ctypes.create_string_buffer(256)
    rc = lib.hc_inflate(bytes(data), len(data), hint, seg_out, seg_in_kb, out.ctypes.data, max_out, ctypes.byref(ol),
                        ctypes.byref(iu), ctypes.byref(sg), ctypes.byref(rd), err, len(err))
    return rc, bytes(out[: ol.value]) if rc == 0 else b"", iu.value, sg.value, rd.value, err.value.decode()


def hc_inflate_batch(streams, caps, seg_out=256 << 20):
    """streams decoded as one batch segment by the host build: [(rc, bytes, in_used, crc)]"""
    lib = hostcheck()
    n = len(streams)
    offs = np.zeros(n, np.uint64)
    sizes = np.array([len(s) for s in streams], np.uint64)
    if n:
        offs[1:] = np.cumsum(sizes)[:-1]
    blob = b"".join(bytes(s) for s in streams)
    caps = np.array(caps, np.uint64)
    rcs = np.zeros(n, np.int32)
    osz, iu = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    crcs = np.zeros(n, np.uint32)
    cap = int(caps.sum()) + 1
    out = np.empty(cap, np.uint8)
    err = ctypes.create_string_buffer(256)
    p = ctypes.c_void_p
    lib.hc_inflate_batch.argtypes = [ctypes.c_char_p, p, p, p, c_int, u64, p, p, p, p, p, u64, ctypes.c_char_p,
                                     ctypes.c_size_t]
    rc = lib.hc_inflate_batch(blob, offs.ctypes.data, sizes.ctypes.data, caps.ctypes.data, n, seg_out, rcs.ctypes.data,
                              osz.ctypes.data, iu.ctypes.data, crcs.ctypes.data, out.ctypes.data, cap, err, len(err))
    if rc != 0:
        raise RuntimeError(err.value.decode())
    res, o = [], 0
    for i in range(n):
        m = int(osz[i])
        res.append((int(rcs[i]), bytes(out[o:o + m]), int(iu[i]), int(crcs[i])))
        o += m
    return res


def _headers(fn, data, start):
    data = bytes(data)
    cap = max(1024, len(data) // 64)
    out = np.zeros(cap, np.uint64)
    fn.restype = u64
    fn.argtypes = [ctypes.c_char_p, u64, u64, ctypes.c_void_p, u64]
    n = fn(data, len(data), start, out.ctypes.data, cap)
    assert n <= cap
    return out[:n].tolist()


def grref():
    """the reference's own rescue library (gpuinflate/rescue.cpp via oracle/rescue_shim.cpp), or None"""
    return _load("grref", os.path.join(ROOT, "oracle", "_ref", "libgrref.so"))


def ref_rescue_headers(data, start):
    """Rescuer::find_headers of the reference (rescue.cpp:597-626), or None without oracle/_ref"""
    lib = grref()
    return None if lib is None else _headers(lib.grref_find_headers, data, start)


def hc_rescue_headers(data, start):
    """the same search by the host build of this project's decoder core"""
    return _headers(hostcheck().hc_rescue_headers, data, start)


def damaged(data, sectors=3, seed=9, size=512):
    """data with `sectors` sectors of `size` zero bytes at seeded sector offsets (never the first
    or the last sector), as the reference's rescue tests damage files (realistic.py:344-350)"""
    rng = np.random.default_rng(seed)
    b = bytearray(data)
    n = len(b) // size
    if n < 3:
        return bytes(b)
    for s in rng.choice(np.arange(1, n - 1), min(sectors, n - 2), replace=False):
        b[s * size:(s + 1) * size] = bytes(size)
    return bytes(b)
