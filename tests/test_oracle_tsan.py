"""ThreadSanitizer run of the oracle's OpenMP render (host code only): the
oracle and a small driver (oracle/tsan_render.c) built by ROCm's clang with
-fsanitize=thread and its OpenMP runtime, whose Archer tool describes the
OpenMP synchronisation to TSan.  The driver renders a small scene (matte,
metal and emissive voxels, pinhole and thin lens, 4 bounces) with several
threads and checks that images, FP64 sums and counters equal the one-thread
render; TSan must report nothing."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"


@pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(ARCHER)), reason="ROCm clang / Archer not present")
@pytest.mark.parametrize("threads", [4, 8])
def test_oracle_openmp_render_has_no_data_race(threads):
    oracle_dir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-C", oracle_dir, "tsan"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ)
    # the OpenMP runtime itself is not instrumented: its internal locking is
    # described by Archer, and its own accesses are ignored
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66 ignore_noninstrumented_modules=1"
    env["OMP_TOOL_LIBRARIES"] = ARCHER
    r = subprocess.run([os.path.join(oracle_dir, "build", "tsan_render"), str(threads)], capture_output=True,
                       text=True, env=env, timeout=600)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert f"tsan_render ok threads {threads}" in r.stdout
