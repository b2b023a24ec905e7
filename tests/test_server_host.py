"""Host-side logic of the build server that needs no GPU: the range
allocator behind its mesh memory (best fit, merging of freed neighbours,
chunk return), fuzzed on the CPU with address and undefined-behaviour
sanitizers."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_arena_fuzz(tmp_path):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "arena_fuzz")
    src = os.path.join(ROOT, "tests", "cpp", "arena_fuzz.cc")
    inc = os.path.join(ROOT, "webgpu-lvbh-demo_amd", "csrc")
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", inc, src, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def test_alloc_table_extents(tmp_path):
    """The extent table behind the C ABI's buffer checks: an upload, download
    or launch over an lbvh_alloc buffer is accepted only when it stays inside
    that buffer (tests/cpp/alloc_table_test.cc, against brute force)."""
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "alloc_table_test")
    src = os.path.join(ROOT, "tests", "cpp", "alloc_table_test.cc")
    inc = os.path.join(ROOT, "webgpu-lvbh-demo_amd", "csrc")
    subprocess.run([gxx, "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", inc, src, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
