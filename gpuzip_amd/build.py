"""Builds the native libraries in-tree (gpuzip_amd/lib/), so they travel with the tree.

    python -m gpuzip_amd.build          # everything
    libgpuinflate.so   the MI355X decoder (hipcc, gfx950): the product, behind include/gpuinflate.h
    libgzrescue.so     the rescue of damaged files (hipcc, gfx950), behind include/gzrescue.h
    libgzhost.so       the gzip member driver that calls it (C++; the reference host's gz.rs role)
    libgzgen.so        seeded synthetic inputs and a parallel zlib compressor (tests, benchmark)
    libgzhostcheck.so  the decoding core and planner built for the CPU (tests only)
    librescuecheck.so  the rescue's decisions over a CPU backend (tests only)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GZ_OFFLOAD_ARCH", "gfx950")
# packed FP32 instructions are switched off for all device code built here
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r


def build_gpuinflate():
    out = os.path.join(LIB, "libgpuinflate.so")
    _run([HIPCC, *HIP_FLAGS, "-I" + INCLUDE, "-o", out, os.path.join(CSRC, "gpuinflate.hip")])
    return out


def build_gzrescue():
    out = os.path.join(LIB, "libgzrescue.so")
    _run([HIPCC, *HIP_FLAGS, "-I" + INCLUDE, "-o", out, os.path.join(CSRC, "gzrescue.hip"), "-L" + LIB, "-lgpuinflate",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_host_libs():
    outs = []
    gen = os.path.join(LIB, "libgzgen.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-o", gen,
          os.path.join(CSRC, "datagen.cpp"), "-lz", "-lpthread"])
    outs.append(gen)
    hc = os.path.join(LIB, "libgzhostcheck.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-o", hc,
          os.path.join(CSRC, "hostcheck.cpp")])
    outs.append(hc)
    rc = os.path.join(LIB, "librescuecheck.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-pthread", "-I" + INCLUDE, "-o", rc,
          os.path.join(CSRC, "rescuecheck.cpp")])
    outs.append(rc)
    gzh_src = os.path.join(CSRC, "gzhost.cpp")
    if os.path.exists(gzh_src):
        gzh = os.path.join(LIB, "libgzhost.so")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-I" + INCLUDE, "-o", gzh,
              gzh_src, "-L" + LIB, "-lgpuinflate", "-Wl,-rpath,$ORIGIN"])
        outs.append(gzh)
    return outs


def build_oracle():
    """The test oracle (oracle/liboracle.so) and, where the reference sources are present,
    oracle/_ref (the reference's own decoder)."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    os.makedirs(LIB, exist_ok=True)
    outs = [build_gpuinflate(), build_gzrescue()]
    outs += build_host_libs()
    build_oracle()
    return outs


if __name__ == "__main__":
    for p in build_all():
        print("built", os.path.relpath(p, ROOT))
