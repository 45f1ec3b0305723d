"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the threaded text reader/writer and binary checkpoints of the library's
tcym_io.cpp and the oracle's sweeps, through tests/sanitize/san_check.cpp,
built here with g++ (-fsanitize=address,undefined, reports abort)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_host_code_under_sanitizers(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle")]
    obj = str(tmp_path / "oracle.o")
    subprocess.run(["gcc", "-std=gnu11", "-fopenmp", "-Wall", "-Wno-unused-function", "-c",
                    os.path.join(ROOT, "oracle", "tcym_oracle.c"), "-o", obj] + san + inc, check=True)
    exe = str(tmp_path / "san_check")
    subprocess.run(["g++", "-std=c++17", "-fopenmp", "-Wall", os.path.join(SAN, "san_check.cpp"),
                    os.path.join(ROOT, "topologicalcymsolver_amd", "csrc", "tcym_io.cpp"), obj, "-o", exe, "-lm",
                    "-lpthread"] + san + inc, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "sanitizer checks passed" in r.stdout
