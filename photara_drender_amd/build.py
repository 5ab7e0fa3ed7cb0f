"""Build recipe of the native parts (run by ``__graft_entry__.build()``).

* ``libpdr.so``            hipcc, gfx950: HIP kernels + the C ABI of ``include/pdr.h``;
* ``_photara_drender*.so`` g++, pybind11: the ``_photara_drender`` Python module
  (numpy API of the reference's nanobind module) over the C++ facade in
  ``include/photara_drender/``, linked against ``libpdr.so``;
* ``libpdr_hostsan.so``     the same library with its host-side C++ built under
  AddressSanitizer + UndefinedBehaviorSanitizer (for the sanitizer tests).

Both land next to this file so they travel with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBPDR = os.path.join(HERE, "libpdr.so")

HIP_SOURCES = ["raster.hip", "interp_texture.hip", "refine.hip", "bake.hip", "mlp.hip", "padding.hip", "pdr_api.cpp",
               "pdr_refine.cpp", "pdr_bake.cpp", "pdr_mlp.cpp", "pdr_padding.cpp", "pdr_noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PDR_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every float expression rounds as written (bit parity with the
# reference's operation order); packed-fp32 ops are disabled for gfx950.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result", "-I", INCLUDE, "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:4]) + " ...")
    return r


def build_libpdr(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIBPDR + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIBPDR)
    return LIBPDR


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the C-ABI
# marshalling (the .cpp sources: argument checks, staging, refine/bake/padding
# drivers), linked with the same kernels.  Each -fsanitize= applies to the host
# compilation only (-Xarch_host); device code is built as in libpdr.so.  Loaded
# by the sanitizer tests with the ASan and UBSan runtimes preloaded.
HOST_SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                  "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]
LIBPDR_HOSTSAN = os.path.join(HERE, "libpdr_hostsan.so")


def build_hostsan(jobs: int = 8) -> str:
    san = os.path.join(BUILD, "hostsan")
    os.makedirs(san, exist_ok=True)
    host_sources = [s for s in HIP_SOURCES if s.endswith(".cpp")]

    def compile_one(src):
        obj = os.path.join(san, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, *HOST_SAN_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, host_sources))
    kernels = [os.path.join(BUILD, os.path.splitext(s)[0] + ".o") for s in HIP_SOURCES if not s.endswith(".cpp")]
    tmp = LIBPDR_HOSTSAN + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, *kernels, "-o", tmp])
    os.replace(tmp, LIBPDR_HOSTSAN)
    return LIBPDR_HOSTSAN


def build_pymodule() -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_photara_drender" + suffix)
    tmp = out + ".tmp"
    _run([
        "g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-fvisibility=hidden",
        "-I", INCLUDE, "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
        os.path.join(CSRC, "bindings.cpp"), "-o", tmp,
        "-L", HERE, "-lpdr", "-Wl,-rpath,$ORIGIN",
    ])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_libpdr()
    build_pymodule()
    build_hostsan()


if __name__ == "__main__":
    build_all()
    print("built", LIBPDR)
