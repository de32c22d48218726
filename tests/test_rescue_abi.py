"""The rescue's drop-in boundary without a GPU: libgzrescue.so loads, exports exactly what
include/gzrescue.h declares (the reference's gr_* interface, ABI 4, rescue.h:38-107, and this
build's gzr_* additions), and answers bad arguments and a missing device with its error codes
before any rescue work."""
import ctypes
import gzip
import os
import re
import subprocess

import rescue_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpuzip_amd", "lib", "libgzrescue.so")
GR_OK, GR_ERR_ARG, GR_ERR_INTERNAL = 0, 3, 6


def declared():
    h = open(os.path.join(ROOT, "include", "gzrescue.h")).read()
    return sorted(set(re.findall(r"GR_API\s+[\w\s\*]+?\b(gr_\w+|gzr_\w+)\s*\(", h)))


def test_exports_exactly_the_header():
    assert os.path.exists(LIB), "libgzrescue.so is not built (python -m gpuzip_amd.build)"
    want = declared()
    assert {"gr_abi_version", "gr_rescue", "gr_rescue_raw", "gr_result_stats", "gr_result_pieces", "gr_result_notes",
            "gr_free"} < set(want)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    syms = sorted(l.split()[-1] for l in out.splitlines() if " T " in l)
    assert [s for s in syms if not s.startswith(("_init", "_fini"))] == want


def test_layouts_version_and_arguments():
    lib, _ = rescue_lib.product()
    lib.gr_abi_version.restype = ctypes.c_uint32
    assert lib.gr_abi_version() == 4
    assert ctypes.sizeof(rescue_lib.Options) == 24
    assert ctypes.sizeof(rescue_lib.Piece) == 72
    assert ctypes.sizeof(rescue_lib.Stats) == 120  # include/gzrescue.h compiled with gcc: 24, 72, 120
    res = ctypes.c_void_p()
    opt = rescue_lib.Options(ctypes.sizeof(rescue_lib.Options), 0, 0, 0, 0, 0, 0)
    assert lib.gr_rescue(None, 5, ctypes.byref(opt), rescue_lib.SINK(lambda *a: 0), None, ctypes.byref(res)) == GR_ERR_ARG
    assert not res.value
    lib.gr_result_stats.restype = ctypes.c_void_p
    assert lib.gr_result_stats(None) is None
    assert lib.gr_result_pieces(None, None) == 0 and lib.gr_result_notes(None, None) == 0
    lib.gr_free(None)
    assert lib.gzr_set_device(-1) == GR_ERR_ARG


def test_no_device_is_an_internal_error():
    """on a device that does not exist the rescue returns GR_ERR_INTERNAL and says why; it never
    falls back to rescuing on the host"""
    lib, _ = rescue_lib.product()
    lib.gzr_last_error.restype = ctypes.c_char_p
    assert lib.gzr_set_device(99) == GR_OK
    try:
        data = gzip.compress(b"a line of text\n" * 1000)
        got = rescue_lib.rescue(rescue_lib.product(), data)
        assert got[0] == GR_ERR_INTERNAL and got[1] == b"" and got[3] is None
        assert b"GPU" in lib.gzr_last_error()
    finally:
        assert lib.gzr_set_device(0) == GR_OK
