"""Build step for the HIP library (libstdgpu_amd.so) and the CPU oracle.

The library is compiled in-tree for gfx950 with hipcc; the oracle (test infrastructure)
with g++ through oracle/Makefile.  `python -m stdgpu_amd.build` runs both.
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "stdgpu_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libstdgpu_amd.so")
ROOFLINE_LIB = os.path.join(PKG, "libstdgpu_amd_roofline.so")
SOURCES = ["htab.hip", "noalloc.hip", "ring.hip", "bitset.hip", "atomic.hip", "server.hip", "kind_set_u32.hip", "kind_set_u64.hip", "kind_map_u64_u32.hip",
           "kind_map_u32_u32.hip", "kind_map_u64_u64.hip"]
HEADERS = [
    os.path.join(ROOT, "include", "stdgpu_amd", "htab.h"),
    os.path.join(ROOT, "include", "stdgpu_amd", "htab_device.hpp"),
    os.path.join(ROOT, "include", "stdgpu_amd", "atomic_device.hpp"),
    os.path.join(ROOT, "include", "stdgpu_amd", "ring_device.hpp"),
    os.path.join(CSRC, "htab_kernels.hpp"),
    os.path.join(CSRC, "wave.hpp"),
    os.path.join(CSRC, "htab_internal.hpp"),
    os.path.join(CSRC, "htab_ops.hpp"),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: no packed FP32 instructions are generated (none are needed here).
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd: list[str], cwd: str | None = None) -> None:
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build_library(out: str = LIB, sources=SOURCES) -> str:
    """Compiles each translation unit (in parallel, one process per unit) and links the library."""
    from concurrent.futures import ThreadPoolExecutor

    build_dir = os.path.join(PKG, "_build")
    os.makedirs(build_dir, exist_ok=True)
    objs = [os.path.join(build_dir, src.replace(".hip", ".o")) for src in sources]
    jobs = max(1, min(len(sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
                for src, obj in zip(sources, objs)]
        for f in futs:
            f.result()
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_roofline() -> str:
    return build_library(ROOFLINE_LIB, ["roofline.hip"])


CPP_TESTS = ["test_facade", "test_atomic", "test_deque", "test_bitset", "test_vector", "test_reference_fixture"]


def build_cpp_tests() -> list[str]:
    """hipcc builds of the C++ facade tests (tests/cpp/*.hip) against the in-tree library."""
    out = []
    src_dir = os.path.join(ROOT, "tests", "cpp")
    bin_dir = os.path.join(src_dir, "_build")
    os.makedirs(bin_dir, exist_ok=True)
    for name in CPP_TESTS:
        exe = os.path.join(bin_dir, name)
        _run([HIPCC, *HIP_FLAGS, "-O2", os.path.join(src_dir, name + ".hip"), "-o", exe,
              "-L", PKG, "-lstdgpu_amd", "-Wl,-rpath,$ORIGIN/../../../stdgpu_amd"])
        out.append(exe)
    return out


def main() -> None:
    print("library:", build_library())
    print("roofline:", build_roofline())
    print("cpp tests:", build_cpp_tests())
    print("oracle :", build_oracle())


if __name__ == "__main__":
    main()
