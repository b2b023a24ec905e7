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


EXTRAS = os.path.join(ROOT, "tests", "js", "extras_cli.js")


def run_extras(*args):
    node = shutil.which("node")
    assert node, "node is required for the JS boundary test"
    print(f"node extras_cli {args[0]}", flush=True)
    r = subprocess.run([node, EXTRAS, *map(str, args)], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr


def test_js_batched_matches_oracle(tmp_path, lb, ora):
    """initBVHBatchBuild: every mesh of the batch equals its own oracle build."""
    sizes = [1, 300, 4097, 0, 20000, 2]
    meshes = [lb.wl_uniform(70 + k, n, 0.01) if n else np.zeros((0, 12), np.float32)
              for k, n in enumerate(sizes)]
    off = np.zeros(len(sizes) + 1, np.uint32)
    off[1:] = np.cumsum(sizes)
    np.concatenate(meshes).astype(np.float32).tofile(tmp_path / "tri.f32")
    off.tofile(tmp_path / "off.u32")
    run_extras("batched", tmp_path / "tri.f32", tmp_path / "off.u32", tmp_path / "b")
    got_bvh = np.fromfile(str(tmp_path / "b.bvh"), np.uint8)
    got_otri = np.fromfile(str(tmp_path / "b.otri"), np.uint32)
    for k, m in enumerate(meshes):
        a, b = int(off[k]), int(off[k + 1])
        if a == b:
            continue
        wb, wo, _, _ = ora.build(m, None, ora.DIV_MUL_RCP)
        assert np.array_equal(got_bvh[64 * a:64 * b], wb), k
        assert np.array_equal(got_otri[12 * a:12 * b], wo.view(np.uint32).ravel()), k


def test_js_raycast_matches_oracle(tmp_path, lb, ora):
    n = 5000
    tri = lb.wl_uniform(71, n, 0.02)
    rng = np.random.default_rng(9)
    rays = np.zeros((2048, 8), np.float32)
    o = rng.uniform(-0.5, 1.5, (2048, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (2048, 3)).astype(np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, t - o
    tri.tofile(tmp_path / "tri.f32")
    rays.tofile(tmp_path / "rays.f32")
    run_extras("raycast", tmp_path / "tri.f32", n, tmp_path / "rays.f32", tmp_path / "r")
    bvh = np.fromfile(str(tmp_path / "r.bvh"), np.uint8)
    otri = np.fromfile(str(tmp_path / "r.otri"), np.float32).reshape(-1, 12)
    hits = np.fromfile(str(tmp_path / "r.hits"), np.float32).reshape(-1, 4)
    assert np.array_equal(hits.view(np.uint32), ora.raycast(bvh, otri, rays).view(np.uint32))
    assert np.array_equal(hits.view(np.uint32),
                          ora.raycast_bruteforce(otri, rays).view(np.uint32))


def test_js_server_build_and_refit(tmp_path, lb, ora):
    """createServer: a host-sourced build equals the oracle build; a jitter
    refit equals the oracle refit of the jittered vertices it produced."""
    n = 30000
    tri = lb.wl_clustered(72, n, 0.05)
    tri.tofile(tmp_path / "tri.f32")
    run_extras("serve", tmp_path / "tri.f32", n, tmp_path / "s")
    wb, wo, wsk, wpm = ora.build(tri, None, ora.DIV_MUL_RCP)
    assert np.array_equal(np.fromfile(str(tmp_path / "s.build.bvh"), np.uint8), wb)
    assert np.array_equal(np.fromfile(str(tmp_path / "s.build.otri"), np.uint32),
                          wo.view(np.uint32).ravel())
    # the refit's O_TRI holds the jittered vertices in leaf order
    jit_otri = np.fromfile(str(tmp_path / "s.refit.otri"), np.float32).reshape(-1, 12)
    jit = np.empty_like(jit_otri)
    jit[wpm] = jit_otri
    rb, ro = ora.refit(jit, wsk, wpm)
    assert np.array_equal(np.fromfile(str(tmp_path / "s.refit.bvh"), np.uint8), rb)
    assert not np.array_equal(jit, tri)


def make_obj(nx=150, ny=100):
    """An OBJ model as a modeller writes one: a height grid of quads (fan
    triangulated by the parser), some faces as triangles with negative
    indices and v/vt/vn references, comments, and a second object whose
    vertices the demo ignores."""
    rng = np.random.default_rng(5)
    lines = ["# test model", "o ground"]
    h = rng.uniform(0.0, 0.3, (ny + 1, nx + 1))
    for j in range(ny + 1):
        for i in range(nx + 1):
            lines.append(f"v {i / nx:.6f} {h[j, i]:.6f} {j / ny * 0.7:.6f}")
    lines += ["vt 0 0", "vn 0 1 0"]
    vid = lambda i, j: j * (nx + 1) + i + 1
    for j in range(ny):
        for i in range(nx):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            if (i + j) % 7 == 0:
                lines.append(f"f {a}/1/1 {b}/1/1 {c}/1/1")
                lines.append(f"f {a} {c} {d}")
            else:
                lines.append(f"f {a} {b} {c} {d}")
    # a pentagon from the last vertices, by negative index
    lines.append("v 0.2 0.9 0.1")
    lines.append("v 0.3 0.95 0.1")
    lines.append("f -1 -2 -5 -7 -9")
    lines += ["o props", "v 5 5 5", "v 6 5 5", "v 5 6 5", "f 1 2 3"]
    return "\n".join(lines) + "\n"


def test_js_obj_model_through_the_drop_in(tmp_path, ora):
    """The demo's flow (main.js:72-88) on the GPU: parseObj of an OBJ text,
    then initBVHBuild(device).build(TRI_ARR, NUM_TRIS, BOUNDS).  BOUNDS
    excludes the two floor triangles, whose keys saturate; the result equals
    the oracle's build of the same f32 triangles and bounds."""
    node = shutil.which("node")
    assert node, "node is required for the JS boundary test"
    obj = tmp_path / "model.obj"
    obj.write_text(make_obj())
    prefix = str(tmp_path / "o")
    print("node obj_build_cli", flush=True)
    r = subprocess.run([node, os.path.join(ROOT, "tests", "js", "obj_build_cli.js"), str(obj),
                        prefix], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    n = json.loads(r.stdout.strip().splitlines()[-1])["NUM_TRIS"]
    assert n == 2 * 150 * 100 + 3 + 2   # quads and split quads, pentagon fan, floor
    tri = np.fromfile(prefix + ".tri", np.float32).reshape(-1, 12)
    bounds = np.fromfile(prefix + ".bounds", np.float32)
    # the floor lies outside the model's bounds
    assert tri[-2:, 0].min() < bounds[0] and tri[-2:, 4].max() > bounds[3]
    wb, wo, wsk, _ = ora.build(tri, bounds, ora.DIV_MUL_RCP)
    assert np.array_equal(np.fromfile(prefix + ".bvh", np.uint8), wb)
    assert np.array_equal(np.fromfile(prefix + ".otri", np.uint32), wo.view(np.uint32).ravel())
