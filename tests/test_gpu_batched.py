"""Batched builds (lbvh_build_batched): each mesh of a batch must come out
byte-identical to a single-mesh build of that mesh, and graph replays must
keep giving the same bytes with other work in between."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_batch(lb, sizes, seed):
    meshes = [lb.wl_uniform(seed + k, int(n), 0.01) if n else np.zeros((0, 12), np.float32)
              for k, n in enumerate(sizes)]
    offsets = np.zeros(len(sizes) + 1, np.uint32)
    offsets[1:] = np.cumsum(sizes)
    tri = np.concatenate(meshes) if offsets[-1] else np.zeros((0, 12), np.float32)
    return meshes, offsets, tri


def run_batch(ctx, tri, offsets, bounds=None):
    total = int(offsets[-1])
    d = ctx.to_device(tri)
    bvh, otri = ctx.alloc(64 * total), ctx.alloc(48 * total)
    sk, pm = ctx.alloc(4 * total), ctx.alloc(4 * total)
    try:
        ctx.build_batched(d, offsets, bounds, bvh, otri, sk, pm)
        return (bvh.read(np.uint8, 64 * total), otri.read(np.float32, 12 * total).reshape(-1, 12),
                sk.read(np.uint32, total), pm.read(np.int32, total))
    finally:
        for b in (d, bvh, otri, sk, pm):
            b.destroy()


@pytest.mark.parametrize("bits", [8, 10])
@pytest.mark.parametrize("with_bounds", [False, True])
def test_batch_meshes_match_oracle(ctx, lb, ora, with_bounds, bits):
    sizes = [1, 2, 0, 3, 255, 256, 257, 4096, 4097, 9000, 1, 33333, 5]
    say(f"batch {len(sizes)} meshes bounds={with_bounds} bits={bits}")
    meshes, offsets, tri = make_batch(lb, sizes, 100)
    bounds = None
    if with_bounds:
        bounds = np.concatenate([lb.wl_bounds(m) if len(m) else np.zeros(6, np.float32)
                                 for m in meshes])
    ctx.set_tuning(lb.TUNE_SORT_BITS, bits)
    try:
        bvh, otri, sk, pm = run_batch(ctx, tri, offsets, bounds)
    finally:
        ctx.set_tuning(lb.TUNE_SORT_BITS, 0)
    for k, m in enumerate(meshes):
        a, b = int(offsets[k]), int(offsets[k + 1])
        if a == b:
            continue
        wb, wo, wsk, wpm = ora.build(m, None if bounds is None else bounds[6 * k:6 * k + 6], 0)
        assert np.array_equal(sk[a:b], wsk), k
        assert np.array_equal(pm[a:b], wpm), k
        assert np.array_equal(bvh[64 * a:64 * b], wb), k
        assert np.array_equal(otri[a:b].view(np.uint32), wo.view(np.uint32)), k


def test_batch_graph_replays(ctx, lb):
    """Replays of the captured graph, with eager work between them, give the
    same bytes as per-mesh lbvh_build calls."""
    rng = np.random.default_rng(3)
    sizes = np.exp(rng.uniform(np.log(256), np.log(70000), 48)).astype(np.uint32)
    meshes, offsets, tri = make_batch(lb, sizes, 500)
    total = int(offsets[-1])
    want = np.concatenate([ctx.build_arrays(m, None)[0] for m in meshes])
    d = ctx.to_device(tri)
    bvh, otri = ctx.alloc(64 * total), ctx.alloc(48 * total)
    other = lb.wl_clustered(7, 50000, 0.05)
    try:
        for rep in range(4):
            say(f"graph replay {rep}")
            ctx.build_batched(d, offsets, None, bvh, otri)
            got = bvh.read(np.uint8, 64 * total)
            assert np.array_equal(got, want), rep
            ctx.build_arrays(other, None)  # eager work between replays
    finally:
        for b in (d, bvh, otri):
            b.destroy()


def test_batch_log_uniform_256(ctx, lb):
    """256 meshes, log-uniform sizes in [256, 262144] (the 2,048-mesh
    benchmark's distribution), against per-mesh builds on the GPU."""
    sizes = lb.wl_log_uniform_sizes(lb.BATCH_SEED, 256, 256, 262144)
    say(f"batch 256 log-uniform meshes, {int(sizes.sum())} triangles")
    meshes, offsets, tri = make_batch(lb, sizes, 1000)
    bvh, otri, sk, pm = run_batch(ctx, tri, offsets)
    for k in range(0, len(meshes), 17):
        a, b = int(offsets[k]), int(offsets[k + 1])
        wb, wo = ctx.build_arrays(meshes[k], None)
        assert np.array_equal(bvh[64 * a:64 * b], wb), k
        assert np.array_equal(otri[a:b], wo), k


def test_batch_benchmark_shape_every_mesh_matches_oracle(lb, ora):
    """The batched benchmark's exact workload: 2,048 uniform meshes, sizes
    log-uniform in [256, 262144] (78.9M triangles), GPU bounds, on a fresh
    untuned context as the benchmark runs it.  Every mesh's sorted keys,
    permutation, BVH nodes and O_TRI equal the oracle's multi-threaded
    build of that mesh alone, and the per-mesh lbvh_build loop the speedup
    is measured against gives the same bytes."""
    meshes = 2048
    sizes = lb.wl_log_uniform_sizes(lb.BATCH_SEED, meshes, 256, 262144)
    offsets = np.zeros(meshes + 1, np.uint32)
    offsets[1:] = np.cumsum(sizes)
    total = int(offsets[-1])
    say(f"batch benchmark shape: {meshes} meshes, {total} triangles")
    tri = np.empty((total, 12), np.float32)
    for k in range(meshes):
        tri[offsets[k]:offsets[k + 1]] = lb.wl_uniform(lb.BATCH_SEED * 100000 + k,
                                                       int(sizes[k]), 0.01)
    with lb.Context(0) as ctx:
        d = ctx.to_device(tri)
        bvh, otri = ctx.alloc(64 * total), ctx.alloc(48 * total)
        sk, pm = ctx.alloc(4 * total), ctx.alloc(4 * total)
        try:
            ctx.build_batched(d, offsets, None, bvh, otri, sk, pm)
            got_bvh = bvh.read(np.uint8, 64 * total)
            got_otri = otri.read(np.uint32, 12 * total).reshape(-1, 12)
            got_sk, got_pm = sk.read(np.uint32, total), pm.read(np.int32, total)
            say("batch benchmark shape: per-mesh loop")
            for k in range(meshes):
                o = int(offsets[k])
                ctx.build(d.ptr + 48 * o, int(sizes[k]), None, bvh.ptr + 64 * o,
                          otri.ptr + 48 * o)
            loop_bvh = bvh.read(np.uint8, 64 * total)
            loop_otri = otri.read(np.uint32, 12 * total).reshape(-1, 12)
        finally:
            for b in (d, bvh, otri, sk, pm):
                b.destroy()
    assert np.array_equal(loop_bvh, got_bvh)
    assert np.array_equal(loop_otri, got_otri)
    del loop_bvh, loop_otri
    say("batch benchmark shape: oracle per mesh")
    for k in range(meshes):
        a, b = int(offsets[k]), int(offsets[k + 1])
        wb, wo, wsk, wpm, _ = ora.build_mt(tri[a:b], None, ora.DIV_MUL_RCP)
        assert np.array_equal(got_sk[a:b], wsk), k
        assert np.array_equal(got_pm[a:b], wpm), k
        assert np.array_equal(got_bvh[64 * a:64 * b], wb), k
        assert np.array_equal(got_otri[a:b], wo.view(np.uint32)), k
