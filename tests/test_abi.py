"""C-ABI checks that need no GPU: the library loads, exports every symbol the
headers declare, and host-side argument handling reports errors instead of
crashing."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbvh_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_lbvh_exports_every_declared_symbol(lb):
    names = declared("lbvh.h")
    assert len(names) >= 20
    syms = exported(lb.LIB_PATH)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    # and the Python binding binds each of them
    assert sorted(lb.SIGNATURES) == names


def test_workloads_exports_every_declared_symbol(lb):
    names = declared("lbvh_workloads.h")
    syms = exported(lb.WL_PATH)
    assert not [n for n in names if n not in syms]
    assert sorted(lb.WL_SIGNATURES) == names


def test_library_links_hip_runtime_only(lb):
    out = subprocess.run(["ldd", lb.LIB_PATH], capture_output=True, text=True).stdout
    assert "libamdhip64" in out
    assert "oracle" not in out and "torch" not in out


def test_status_strings(lb):
    L = lb.load_library()
    for code, text in lb.STATUS.items():
        assert L.lbvh_status_string(code).decode() == text


def test_null_context_is_rejected(lb):
    L = lb.load_library()
    assert L.lbvh_build(None, None, 0, None, None, None, None, None) == 1
    assert L.lbvh_synchronize(None) == 1
    assert L.lbvh_last_error(None) == b"null context"
    assert L.lbvh_create(0, 0, None) == 1


def test_create_reports_missing_device(lb):
    """Without a GPU (this container) creation fails with a status code."""
    L = lb.load_library()
    h = C.c_void_p()
    st = L.lbvh_create(10 ** 6, 0, C.byref(h))
    assert st == 6 and not h.value


def test_node_boundary_loads_addon(lb):
    """The Node.js module loads the N-API addon, exports the reference's entry
    points, rounds host inputs like the reference and reports a missing
    device as an Error carrying the status code (tests/js/boundary_cpu.js)."""
    if not os.path.exists(os.path.join(lb.LIB_DIR, "lbvh.node")):
        pytest.fail("lbvh.node was not built (make -C webgpu-lvbh-demo_amd addon)")
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    r = subprocess.run([node, os.path.join(ROOT, "tests", "js", "boundary_cpu.js")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "boundary_cpu ok" in r.stdout
