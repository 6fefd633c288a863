"""Memory and undefined-behaviour checks of the host code, on the CPU.

Two programs are built with AddressSanitizer (leak detection included) and
UndefinedBehaviorSanitizer, every report fatal:
  * tests/sanitize/oracle_driver.c with the oracle's C sources: every oracle
    function on small and degenerate inputs (n = 0, 1, 2; k = 0; zeros,
    duplicates and NaN);
  * tests/sanitize/capi_driver.c with the C-ABI layer (gt_capi.cpp, its host
    code instrumented) and the library's kernel objects: the argument checks
    of every entry point, which reject before any device call.  The HIP
    runtime itself is not instrumented, so leak detection is off for it.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OBJ = os.path.join(ROOT, "gputools_amd", "csrc", "_build")
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    assert r.returncode == 0, "%s\n%s\n%s" % (" ".join(cmd), r.stdout[-4000:], r.stderr[-4000:])
    return r


def test_oracle_under_asan_and_ubsan(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path / "oracle_driver")
    srcs = [os.path.join(ROOT, "oracle", f) for f in ("dist_oracle.c", "hclust_oracle.c", "cor_oracle.c")]
    _run([gcc, "-O1", "-g", "-std=c99", "-fopenmp", "-fno-omit-frame-pointer",
          "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
          os.path.join(SAN, "oracle_driver.c")] + srcs + ["-lm", "-o", exe])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = _run([exe], env=env)
    assert "oracle sanitizer run: ok" in r.stdout


def test_capi_checks_under_asan_and_ubsan(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    objs = sorted(os.path.join(OBJ, f) for f in os.listdir(OBJ)
                  if f.endswith(".o") and not f.startswith("gt_capi")) if os.path.isdir(OBJ) else []
    if not objs:
        pytest.skip("library objects not built")
    capi = str(tmp_path / "gt_capi_san.o")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=all"]
    _run([HIPCC, "-O1", "-g", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          "--offload-arch=gfx950", "-mcode-object-version=5", "-x", "hip"] + san +
         ["-c", os.path.join(ROOT, "gputools_amd", "csrc", "gt_capi.cpp"), "-o", capi])
    exe = str(tmp_path / "capi_driver")
    _run([HIPCC, "--offload-arch=gfx950", "-mcode-object-version=5", "-g",
          "-fsanitize=address,undefined", "-x", "c", os.path.join(SAN, "capi_driver.c"), "-x", "none",
          capi] + objs + ["-ldl", "-o", exe])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = _run([exe], env=env)
    assert "C-ABI sanitizer run: ok" in r.stdout
