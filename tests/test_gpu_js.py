"""The Node.js drop-in (bvh/index.js over the N-API addon) on the GPU,
against the CPU oracle: initBVHBuild(device).build(TRI_ARRAY, N, BOUNDS) with
a plain JS Array as main.js:88 passes it, and the five per-stage execute()
shims in build-bvh.js order."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "tests", "js", "build_cli.js")


def run_cli(tmp_path, tri, bounds, mode):
    node = shutil.which("node")
    assert node, "node is required for the JS boundary test"
    tri_path = tmp_path / "tri.f32"
    np.ascontiguousarray(tri, np.float32).tofile(tri_path)
    barg = "-"
    if bounds is not None:
        barg = str(tmp_path / "bounds.json")
        with open(barg, "w") as f:
            json.dump({"min": [float(x) for x in bounds[:3]],
                       "max": [float(x) for x in bounds[3:]]}, f)
    prefix = str(tmp_path / "out")
    print(f"node build_cli {mode} n={len(tri)}", flush=True)
    r = subprocess.run([node, CLI, str(tri_path), str(len(tri)), barg, prefix, mode],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    return prefix


@pytest.mark.parametrize("mode", ["build", "stages"])
def test_js_entry_points_match_oracle(tmp_path, lb, ora, mode):
    tri = lb.workload("uniform-64k")
    bounds = lb.wl_bounds(tri)
    prefix = run_cli(tmp_path, tri, bounds, mode)
    wb, wo, wsk, wpm = ora.build(tri, bounds, ora.DIV_MUL_RCP)
    assert np.array_equal(np.fromfile(prefix + ".bvh", np.uint8), wb)
    assert np.array_equal(np.fromfile(prefix + ".otri", np.uint32), wo.view(np.uint32).ravel())
    if mode == "stages":
        assert np.array_equal(np.fromfile(prefix + ".keys", np.uint32), wsk)
        assert np.array_equal(np.fromfile(prefix + ".idx", np.int32), wpm)
        assert np.array_equal(np.fromfile(prefix + ".par", np.int32), ora.radix_tree(wsk))


def test_js_build_without_bounds(tmp_path, lb, ora):
    tri = lb.wl_clustered(60, 30000, 0.05)
    prefix = run_cli(tmp_path, tri, None, "build")
    wb, wo, _, _ = ora.build(tri, None, ora.DIV_MUL_RCP)
    assert np.array_equal(np.fromfile(prefix + ".bvh", np.uint8), wb)
