"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU.

* The oracle (oracle/pdr_oracle.cpp) built with g++ -fsanitize=address,undefined
  (``make -C oracle sanitize``) runs the CPU oracle suites - the reference's
  golden vectors, the closed-form pins, padding, bake and atlas checks - in a
  subprocess with libasan preloaded.
* The C-ABI library with its host-side C++ built under the same sanitizers
  (libpdr_hostsan.so: -Xarch_host -fsanitize=..., kernels unchanged) runs
  tests/hostsan_driver.py's GPU-free paths with the ASan and UBSan runtimes
  preloaded (the GPU paths run in tests/test_gpu_sanitizers.py).  The runtimes
  are gcc's: clang's ASan runtime also intercepts the HSA device-memory
  allocator, which it cannot serve without XNACK.
* Each sanitizer is shown to be live: a planted out-of-bounds read (a host
  array one row shorter than the count passed) is reported.
Any report fails the run (UBSan built with -fno-sanitize-recover).
"""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_ASAN = os.path.join(ROOT, "oracle", "_build", "libpdr_oracle_asan.so")
HOSTSAN = os.path.join(ROOT, "photara_drender_amd", "libpdr_hostsan.so")
ORACLE_SUITES = ["tests/test_oracle.py", "tests/test_oracle_pins.py", "tests/test_golden.py", "tests/test_padding.py",
                 "tests/test_bake.py", "tests/test_atlas.py"]


def gcc_asan():
    return subprocess.check_output(["g++", "-print-file-name=libasan.so"], text=True).strip()


def host_runtimes():
    """gcc's ASan + UBSan runtimes, for the library whose host code clang instrumented."""
    return " ".join(subprocess.check_output(["g++", f"-print-file-name={n}"], text=True).strip()
                    for n in ("libasan.so", "libubsan.so"))


def run(preload, extra_env, args, timeout=600):
    env = dict(os.environ, LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1",
               PYTHONPATH=ROOT, **extra_env)
    return subprocess.run([sys.executable, *args], env=env, cwd=ROOT, capture_output=True, text=True,
                          timeout=timeout)


def clean(r):
    text = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, text
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, text


def oracle_asan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    return ORACLE_ASAN


def test_oracle_suites_under_sanitizers():
    r = run(gcc_asan(), {"PDR_ORACLE_LIB": oracle_asan()},
            ["-m", "pytest", *ORACLE_SUITES, "-m", "not gpu", "-q", "-p", "no:cacheprovider"])
    clean(r)
    assert " passed" in r.stdout


PLANTED_ORACLE = """
import ctypes, numpy as np, oracle
F = ctypes.POINTER(ctypes.c_float)
pos = np.ones((200, 3), np.float32)
out = np.empty((201, 4), np.float32)
m = np.eye(4, dtype=np.float32)
oracle.lib().orc_transform_positions(pos.ctypes.data_as(F), 201, m.ctypes.data_as(F), out.ctypes.data_as(F))
"""


def test_oracle_sanitizer_reports_a_planted_overread():
    r = run(gcc_asan(), {"PDR_ORACLE_LIB": oracle_asan()}, ["-c", PLANTED_ORACLE], timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]
    assert "orc_transform_positions" in r.stderr


def test_host_library_cpu_paths_under_sanitizers():
    assert os.path.exists(HOSTSAN), "libpdr_hostsan.so is missing: run __graft_entry__.build()"
    r = run(host_runtimes(), {"PDR_LIB": HOSTSAN}, ["tests/hostsan_driver.py", "cpu"], timeout=300)
    clean(r)
    assert "hostsan driver: ok" in r.stdout


def test_host_library_sanitizer_reports_a_planted_overread():
    assert os.path.exists(HOSTSAN), "libpdr_hostsan.so is missing: run __graft_entry__.build()"
    r = run(host_runtimes(), {"PDR_LIB": HOSTSAN}, ["tests/hostsan_driver.py", "planted"], timeout=120)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr, r.stderr[-2000:]
    assert "pdr_bake.cpp" in r.stderr  # the scale helper behind pdr_suggest_softmax_scale


def test_sanitized_builds_compute_the_same_oracle_values():
    """The sanitizer build is the same source: spot-check one function's bits."""
    code = ("import numpy as np, oracle\n"
            "from photara_drender_amd import scenes\n"
            "s = scenes.chain_scene(level=2, res=(24, 32), tex_size=8, seed=4)\n"
            "r, d, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)\n"
            "np.save(__import__('sys').argv[1], np.concatenate([r.ravel(), d.ravel()]))\n")
    out = os.path.join(ROOT, "oracle", "_build", "asan_spot.npy")
    r = run(gcc_asan(), {"PDR_ORACLE_LIB": oracle_asan()}, ["-c", code, out], timeout=120)
    clean(r)
    import oracle
    from photara_drender_amd import scenes

    s = scenes.chain_scene(level=2, res=(24, 32), tex_size=8, seed=4)
    ref_r, ref_d, _ = oracle.rasterize_forward(s.clip, s.indices, s.resolution)
    got = np.load(out)
    np.testing.assert_array_equal(got.view(np.uint32), np.concatenate([ref_r.ravel(), ref_d.ravel()]).view(np.uint32))
