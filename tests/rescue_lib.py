"""ctypes access to the three rescues for the tests: the reference's own (gpuinflate/rescue.cpp
built from its sources into oracle/_ref/libgrref.so; absent on a box without the reference), the
host build of this project's (gpuzip_amd/lib/librescuecheck.so: rescue_core.h over a CPU backend)
and the product (gpuzip_amd/lib/libgzrescue.so: the same decisions over the GPU). Each gives
(rc, output bytes, mask bytes, stats, pieces, notes). Test infrastructure only."""
import ctypes
import hashlib
import os

from oracle_lib import OrResult, OrStart

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
u8, u32, u64, i32, i64 = ctypes.c_uint8, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int64


class Options(ctypes.Structure):
    _fields_ = [("size", u32), ("fill", u8), ("no_infer", u8), ("cr_lost", u8), ("crs_expected", u8),
                ("max_output", u64), ("cr_budget", u64)]


class Piece(ctypes.Structure):
    _fields_ = [("out_offset", u64), ("length", u64), ("decoded", u64), ("known", u64), ("in_byte_start", u64),
                ("in_byte_end", u64), ("suspect", i32), ("status", i32), ("blocks", u32), ("symbols", u32),
                ("why", ctypes.c_char_p)]


class Stats(ctypes.Structure):
    _fields_ = [("input_size", u64), ("output_size", u64), ("exact_bytes", u64), ("suspect_bytes", u64),
                ("unknown_bytes", u64), ("headers", u32), ("pieces", u32), ("damaged", u32), ("members", u32),
                ("crc_check", i32), ("length_off", i64), ("truncated", i32), ("header_damaged", i32),
                ("input_end", u64), ("inferred_bytes", u64), ("ascii_undone", i32), ("cr_restored", u32),
                ("cr_decoded", u64), ("hidden_damage", i32)]


SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(u8), ctypes.POINTER(u8), ctypes.c_size_t)

_libs = {}


def _bind(path, prefix):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        _libs[path] = None
        return None
    lib = ctypes.CDLL(path)
    f = lambda n: getattr(lib, prefix + n)
    f("gr_rescue").argtypes = [ctypes.c_char_p, u64, ctypes.POINTER(Options), SINK, ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    f("gr_rescue_raw").argtypes = [ctypes.c_char_p, u64, u64, ctypes.c_int, u32, ctypes.POINTER(Options), SINK,
                                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    f("gr_result_stats").restype = ctypes.POINTER(Stats)
    f("gr_result_stats").argtypes = [ctypes.c_void_p]
    f("gr_result_pieces").restype = u32
    f("gr_result_pieces").argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(Piece))]
    f("gr_result_notes").restype = u32
    f("gr_result_notes").argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p))]
    f("gr_free").argtypes = [ctypes.c_void_p]
    if hasattr(lib, prefix + "gzr_result_phases"):
        f("gzr_result_phases").restype = u32
        f("gzr_result_phases").argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double), u32]
    _libs[path] = (lib, prefix)
    return _libs[path]


def reference():
    return _bind(os.path.join(ROOT, "oracle", "_ref", "libgrref.so"), "")


def hostbuild():
    return _bind(os.path.join(ROOT, "gpuzip_amd", "lib", "librescuecheck.so"), "rc_")


def product():
    return _bind(os.path.join(ROOT, "gpuzip_amd", "lib", "libgzrescue.so"), "")


def rescue(which, data, raw=None, fill=0, no_infer=0, max_output=0, cr_lost=0, phases=None):
    """which: a bound library; raw: None (gzip) or (expected_size or None, crc or None); phases: a
    dict that receives the seconds of each step, where the library reports them"""
    lib, pre = which
    f = lambda n: getattr(lib, pre + n)
    out, mask = bytearray(), bytearray()

    def sink(user, d, m, n):
        out.extend(ctypes.string_at(d, n))
        mask.extend(ctypes.string_at(m, n))
        return 0

    cb = SINK(sink)
    opt = Options(ctypes.sizeof(Options), fill, no_infer, cr_lost, 0, max_output, 0)
    res = ctypes.c_void_p()
    data = bytes(data)
    if raw is None:
        rc = f("gr_rescue")(data, len(data), ctypes.byref(opt), cb, None, ctypes.byref(res))
    else:
        size, crc = raw
        rc = f("gr_rescue_raw")(data, len(data), (1 << 64) - 1 if size is None else size, 0 if crc is None else 1,
                                crc or 0, ctypes.byref(opt), cb, None, ctypes.byref(res))
    stats, pieces, notes = None, [], []
    if res.value:
        s = f("gr_result_stats")(res).contents
        stats = {k: getattr(s, k) for k, _ in Stats._fields_}
        pp = ctypes.POINTER(Piece)()
        n = f("gr_result_pieces")(res, ctypes.byref(pp))
        for i in range(n):
            p = pp[i]
            pieces.append(tuple(getattr(p, k) for k, _ in Piece._fields_))
        nn = ctypes.POINTER(ctypes.c_char_p)()
        k = f("gr_result_notes")(res, ctypes.byref(nn))
        notes = [nn[i].decode() for i in range(k)]
        if phases is not None and hasattr(lib, pre + "gzr_result_phases"):
            names, secs = (ctypes.c_char_p * 32)(), (ctypes.c_double * 32)()
            for i in range(min(32, f("gzr_result_phases")(res, names, secs, 32))):
                phases[names[i].decode()] = secs[i]
        f("gr_free")(res)
    return rc, bytes(out), bytes(mask), stats, pieces, notes


def rc_lib():
    path = os.path.join(ROOT, "gpuzip_amd", "lib", "librescuecheck.so")
    lib = ctypes.CDLL(path)
    lib.rc_run.restype = None
    lib.rc_run.argtypes = [ctypes.c_char_p, u64, ctypes.POINTER(OrStart), ctypes.POINTER(OrStart), ctypes.c_int,
                           ctypes.c_void_p, u32, ctypes.c_int, u64, ctypes.POINTER(OrResult)]
    lib.rc_probe.restype = None
    lib.rc_probe.argtypes = [ctypes.c_char_p, u64, u8, u64, u64, u64, ctypes.POINTER(OrStart)]
    return lib


def summary(result):
    """what the golden fixtures keep of a rescue: the code, SHA-256 of the output and the mask, the
    statistics, the pieces and the notes"""
    rc, out, mask, stats, pieces, notes = result
    return {"rc": rc, "out_len": len(out), "out_sha256": hashlib.sha256(out).hexdigest(),
            "mask_sha256": hashlib.sha256(mask).hexdigest(), "stats": stats,
            "pieces": [[x.decode() if isinstance(x, bytes) else x for x in p] for p in pieces], "notes": notes}


def assert_same(got, want):
    """two rescues' results field by field (the output and mask compared without printing them)"""
    assert got[0] == want[0], "return codes differ"
    assert len(got[1]) == len(want[1]), "output lengths differ"
    assert got[1] == want[1], "output bytes differ"
    assert got[2] == want[2], "masks differ"
    assert got[3] == want[3], {k: (got[3].get(k), want[3].get(k)) for k in set(got[3]) | set(want[3])
                               if got[3].get(k) != want[3].get(k)}
    assert len(got[4]) == len(want[4]), ("pieces", len(got[4]), len(want[4]))
    assert got[4] == want[4], next((i, a, b) for i, (a, b) in enumerate(zip(got[4], want[4])) if a != b)
    assert got[5] == want[5], ("notes", got[5], want[5])
