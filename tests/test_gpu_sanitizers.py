"""The C-ABI library's host-side C++ under AddressSanitizer + UndefinedBehavior-
Sanitizer on the GPU: libpdr_hostsan.so (host code built with -Xarch_host
-fsanitize=address,undefined; kernels as in libpdr.so) drives small workloads
through every host-array path of tests/hostsan_driver.py - staging, batched
views, argument errors, the autotune cache text, timing, refine (graph and
eager, seams, TV, checkpoints), bake, bake features and padding - in a
subprocess with gcc's ASan and UBSan runtimes preloaded (clang's ASan runtime
also intercepts the HSA device-memory allocator, which it cannot serve
without XNACK).  Any sanitizer report fails.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTSAN = os.path.join(ROOT, "photara_drender_amd", "libpdr_hostsan.so")


def test_host_library_gpu_paths_under_sanitizers():
    assert os.path.exists(HOSTSAN), "libpdr_hostsan.so is missing: run __graft_entry__.build()"
    runtimes = " ".join(subprocess.check_output(["g++", f"-print-file-name={n}"], text=True).strip()
                        for n in ("libasan.so", "libubsan.so"))
    env = dict(os.environ, LD_PRELOAD=runtimes, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1",
               PDR_LIB=HOSTSAN, PYTHONPATH=ROOT)
    print("hostsan driver: gpu", flush=True)
    r = subprocess.run([sys.executable, "tests/hostsan_driver.py", "gpu"], env=env, cwd=ROOT, capture_output=True,
                       text=True, timeout=240)
    text = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, text
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, text
    assert "hostsan driver: ok" in r.stdout, text
