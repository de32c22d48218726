"""The drop-in boundary without a GPU: libgpuinflate.so loads, exports exactly the symbols
include/gpuinflate.h declares (as the reference's CI checks with nm, .github/workflows/ci.yml:111),
keeps the v1 struct layouts, and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import pytest

from gpuzip_amd import gpuinflate as gi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    h = open(os.path.join(ROOT, "include", "gpuinflate.h")).read()
    return sorted(set(re.findall(r"GI_API\s+[\w\s\*]+?\b(gi_\w+)\s*\(", h)))


def test_header_declares_the_bindings():
    assert declared() == sorted(gi.EXPORTS)


def test_exports_exactly_the_header():
    path = gi.library_path()
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    syms = sorted(l.split()[-1] for l in out.splitlines() if re.search(r" T gi_", l))
    assert syms == declared()
    # nothing else escapes -fvisibility=hidden but the C++ runtime's own
    others = [l.split()[-1] for l in out.splitlines() if " T " in l and " T gi_" not in l]
    assert not [s for s in others if not s.startswith(("_init", "_fini"))], others


def test_layouts_and_version():
    lib = gi.load_library()
    assert lib.gi_abi_version() == 1
    assert ctypes.sizeof(gi.Config) == 16
    assert ctypes.sizeof(gi.Stats) == 80


def test_bad_arguments_before_any_device_call():
    lib = gi.load_library()
    err = ctypes.create_string_buffer(128)
    cfg = gi.Config(8, 0, 0, 0)  # size smaller than sizeof(gi_config)
    ctx = ctypes.c_void_p()
    assert lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, 128) == gi.GI_ERR_ARG
    assert b"too small" in err.value
    assert lib.gi_create(ctypes.byref(gi.Config(16, 0, 0, 0)), None, err, 128) == gi.GI_ERR_ARG
    st = gi.Stats()
    assert lib.gi_inflate(None, b"x", 1, 0, gi.SINK(lambda *a: 0), None, ctypes.byref(st)) == gi.GI_ERR_ARG
    assert lib.gi_last_error(None) == b"no context"
    assert lib.gi_device_name(None) == b""
    lib.gi_destroy(None)


def test_no_device_is_a_gpu_error():
    """without a usable GPU gi_create returns GI_ERR_GPU and a message (the caller then decodes
    on the CPU, gz.rs:231)"""
    lib = gi.load_library()
    cfg = gi.Config(16, 99, 0, 0)  # no such device ordinal on any box
    ctx = ctypes.c_void_p()
    err = ctypes.create_string_buffer(128)
    assert lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, 128) == gi.GI_ERR_GPU
    assert err.value


def test_gzhost_exports_exactly_its_header():
    """libgzhost.so (the member driver above gi_inflate) exports what include/gzhost.h declares"""
    from gpuzip_amd import gz
    h = open(os.path.join(ROOT, "include", "gzhost.h")).read()
    want = sorted(set(re.findall(r"GZ_API\s+[\w\s\*]+?\b(gz_\w+)\s*\(", h)))
    assert want == ["gz_decode", "gz_header"]
    gz.load_library()
    path = os.path.join(os.path.dirname(gi.library_path()), gz.LIB_NAME)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    assert sorted(l.split()[-1] for l in out.splitlines() if re.search(r" T gz_", l)) == want
    assert ctypes.sizeof(gz.Report) == 8 + 8 + 8 + 80 + 3 * 256


def test_tuned_config_checked_before_any_device_call():
    """gi_config_tuned: the reference's 16 bytes, then the schedule; the size field says which
    (gpuinflate.cu:888-892 versions gi_config by it). Fields outside their ranges are
    GI_ERR_ARG, before any device call; in range, creation goes on to the device."""
    lib = gi.load_library()
    assert ctypes.sizeof(gi.ConfigTuned) == 40
    err = ctypes.create_string_buffer(128)
    ctx = ctypes.c_void_p()
    for field, (lo, hi) in gi.TUNING_RANGES.items():
        for v in (lo - 1, hi + 1):
            if v == 0:  # 0 keeps the field's default
                continue
            cfg = gi.make_config(99, tuning={field: v})
            assert lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, 128) == gi.GI_ERR_ARG, (field, v)
            assert field.encode() in err.value
            b = ctypes.c_void_p()
            assert lib.gi_batch_create(ctypes.byref(cfg), 1, ctypes.byref(b), err, 128) == gi.GI_ERR_ARG
        cfg = gi.make_config(99, tuning={field: hi})
        assert lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, 128) == gi.GI_ERR_GPU
    for v in (17, 24, 48):  # lanes_per_group is 16, 32 or 64
        cfg = gi.make_config(99, tuning={"lanes_per_group": v})
        assert lib.gi_create(ctypes.byref(cfg), ctypes.byref(ctx), err, 128) == gi.GI_ERR_ARG, v
    # a config of the 32 bytes before lanes_per_group is read as such, whatever lies after it
    raw = (ctypes.c_uint32 * 10)(32, 99, 0, 0, 4, 0, 0, 0, 17, 1)
    assert lib.gi_create(ctypes.cast(raw, ctypes.c_void_p), ctypes.byref(ctx), err, 128) == gi.GI_ERR_GPU
    # a 16-byte config never reads past its end, whatever lies there
    raw = (ctypes.c_uint32 * 8)(16, 99, 0, 0, 1, 1, 1, 1)
    assert lib.gi_create(ctypes.cast(raw, ctypes.c_void_p), ctypes.byref(ctx), err, 128) == gi.GI_ERR_GPU
    with pytest.raises(ValueError):
        gi.make_config(0, tuning={"wavefronts": 2})


def test_batch_and_segment_limits_checked_before_any_device_call():
    """gi_batch_create takes 1..4 workers (each a context with two HIP streams; a process has four
    hardware queues, so four contexts already share them two streams a queue), and
    gi_config.segment_in_kb is 0 or 16..1048576 KiB: both refused with GI_ERR_ARG before any
    device call"""
    lib = gi.load_library()
    err = ctypes.create_string_buffer(128)
    b = ctypes.c_void_p()
    cfg = gi.make_config(99)
    for workers in (0, 5, 64):
        assert lib.gi_batch_create(ctypes.byref(cfg), workers, ctypes.byref(b), err, 128) == gi.GI_ERR_ARG, workers
        assert b"workers must be 1..4" in err.value
    for workers in (1, 4):  # goes on to the device
        assert lib.gi_batch_create(ctypes.byref(cfg), workers, ctypes.byref(b), err, 128) == gi.GI_ERR_GPU
    ctx = ctypes.c_void_p()
    for kb in (1, 15, (1 << 20) + 1):
        c = gi.Config(16, 99, 0, kb)
        assert lib.gi_create(ctypes.byref(c), ctypes.byref(ctx), err, 128) == gi.GI_ERR_ARG, kb
        assert b"segment_in_kb" in err.value
    for kb in (16, 1 << 20):
        c = gi.Config(16, 99, 0, kb)
        assert lib.gi_create(ctypes.byref(c), ctypes.byref(ctx), err, 128) == gi.GI_ERR_GPU, kb


def test_serve_checked_before_any_device_call():
    """gi_serve_create: a budget under 1 MiB is GI_ERR_ARG before any device call; with no such
    device it is GI_ERR_GPU; gi_serve_submit without a server or request is GI_ERR_ARG"""
    lib = gi.load_library()
    err = ctypes.create_string_buffer(128)
    s = ctypes.c_void_p()
    cfg = gi.make_config(99)
    assert lib.gi_serve_create(ctypes.byref(cfg), 2, 1000, ctypes.byref(s), err, 128) == gi.GI_ERR_ARG
    assert b"budget" in err.value
    assert lib.gi_serve_create(ctypes.byref(cfg), 2, 1 << 30, ctypes.byref(s), err, 128) == gi.GI_ERR_GPU
    assert not s.value
    assert lib.gi_serve_submit(None, None) == gi.GI_ERR_ARG
    assert not lib.gi_serve_wait(None, 0.0)
    lib.gi_serve_destroy(None)
