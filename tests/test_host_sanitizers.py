"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/cpp/host_sanitize.cpp drives the library's host sources (the NetCDF
workspace reader and writer, the flat-city builder) and the CPU oracle,
compiled by g++ with -fsanitize=address,undefined from the source files
themselves (no GPU code is built; the HIP headers are only read for the
__host__ __device__ helpers the builder shares with its device twin).  The
reader is fed truncated and corrupted headers, a header whose variable names
a dimension id out of range, huge counts, offsets past the end of the file,
seeded random byte flips and non-integral icellflag values; every call must
return a status with no sanitizer report.

ThreadSanitizer is not run: the oracle's only threading is OpenMP, and GCC's
libgomp is not built with TSan instrumentation, so TSan reports its own
barriers as races; the oracle's thread-count independence is checked instead
(tests/test_oracle.py, bitwise equal results for 1 and 8 threads)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_workspace import problem, scipy_workspace

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "qes-public_amd", "csrc")


def _have_asan():
    if not shutil.which("g++"):
        return False
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    return r.returncode == 0 and os.path.isabs(r.stdout.strip())


@pytest.mark.skipif(not _have_asan(), reason="g++ with libasan is not available")
def test_host_code_is_clean_under_asan_and_ubsan(qesmc, oracle, tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
    objs = []
    for src in ("qesmc_workspace.cpp", "qesmc_city.cpp"):
        o = str(tmp_path / (src + ".o"))
        subprocess.run(["g++", "-std=c++17", *san, *inc, "-c", os.path.join(CSRC, src), "-o", o], check=True)
        objs.append(o)
    o = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", "-std=gnu11", "-fopenmp", "-ffp-contract=off", *san, "-c",
                    os.path.join(ROOT, "oracle", "qesmc_oracle.c"), "-o", o], check=True)
    objs.append(o)
    exe = str(tmp_path / "host_sanitize")
    subprocess.run(["g++", "-std=c++17", *san, *inc, os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"),
                    *objs, "-fopenmp", "-lm", "-o", exe], check=True)
    # files made here that the reader must refuse: icellflag NaN / beyond int
    p, _, _, _ = problem(oracle, qesmc)
    extra = []
    for q, (t, fill) in enumerate((("f", float("nan")), ("d", 1e20))):
        path = str(tmp_path / f"flag{q}.nc")
        scipy_workspace(path, p, 2, flag_type=t, flag_fill=fill)
        extra.append(path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    out = subprocess.run([exe, str(tmp_path), *extra], capture_output=True, text=True, timeout=600, env=env)
    print(out.stdout[-2000:], out.stderr[-4000:])
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-4000:]
    assert "PASS" in out.stdout
    assert "runtime error" not in out.stderr and "AddressSanitizer" not in out.stderr
