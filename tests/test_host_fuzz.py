"""Randomised robustness run of the library's host code (tests/fuzz/host_fuzz.cpp)
under AddressSanitizer and UndefinedBehaviorSanitizer: the DTA / MGF / FASTA
parsers, the index reader on damaged files, the formatters, digestion and the
modification table must return on any input without a memory error, undefined
behaviour or a runaway loop.  Host code only, built with g++; no GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCES = [os.path.join(ROOT, "tests", "fuzz", "host_fuzz.cpp"),
           os.path.join(ROOT, "hfx_amd", "csrc", "hfx_io.cpp"),
           os.path.join(ROOT, "hfx_amd", "csrc", "hfx_host.cpp")]


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    if not os.path.exists("/usr/include/zlib.h"):
        pytest.skip("zlib headers not found")
    exe = str(tmp_path_factory.mktemp("fuzz") / "host_fuzz")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-I" + os.path.join(ROOT, "include"),
           *SOURCES, "-o", exe, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if r.returncode != 0 and "sanitize" in r.stderr and "cannot find" in r.stderr:
        pytest.skip("sanitizer runtime not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_host_code_under_sanitizers(fuzzer, tmp_path, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([fuzzer, "4000", str(seed), str(tmp_path / "scratch.index")],
                       capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0 and "host_fuzz: 4000 iterations" in r.stdout, \
        (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
