"""GPU parity: every C-ABI entry point against the CPU oracle, bit-exact.

Integer outputs (keys, permutation, parents, child ids) and the fp32 boxes
(exact min/max) must match to the bit; the BVH is compared in the canonical
child order both sides write.  Sizes stay where the oracle finishes in
seconds, plus one full-size (16,777,216-triangle) comparison.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def sort_keys(rng, n, bits=30, dup=0.3):
    keys = rng.integers(0, 2 ** bits, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 1:
        m = rng.random(n) < dup
        keys[m] = keys[rng.integers(0, n, size=int(m.sum()))]
    return keys


def mesh_cases(lb):
    rng = np.random.default_rng(7)
    zeros = rng.uniform(-1, 1, size=(3000, 12)).astype(np.float32)
    zeros[:, 3::4] = np.float32(3.1415)
    mask = rng.random((3000, 12)) < 0.2
    mask[:, 3::4] = False
    zeros[mask] = 0.0
    same = np.tile(lb.wl_uniform(9, 1, 0.01), (777, 1))  # all keys equal
    return {
        "n1": lb.wl_uniform(21, 1, 0.01),
        "n2": lb.wl_uniform(22, 2, 0.01),
        "n3": lb.wl_uniform(23, 3, 0.01),
        "n5": lb.wl_uniform(24, 5, 0.01),
        "n255": lb.wl_uniform(25, 255, 0.01),
        "n256": lb.wl_uniform(26, 256, 0.01),
        "n257": lb.wl_uniform(27, 257, 0.01),
        "n4097": lb.wl_clustered(28, 4097, 0.3),
        "zeros3000": zeros,
        "allsame777": same,
        "clustered100k": lb.wl_clustered(29, 100003, 0.05),
        "uniform-64k": lb.workload("uniform-64k"),
    }


# ---------------------------------------------------------------- stages

@pytest.mark.parametrize("n", [1, 2, 1000, 65536, 1000003])
def test_bounds(ctx, lb, ora, n):
    say(f"bounds n={n}")
    tri = lb.wl_clustered(n, n, 0.05)
    d = ctx.to_device(tri)
    out = ctx.alloc(24)
    ctx.bounds(d, n, out)
    got = out.read(np.float32, 6)
    assert np.array_equal(got, ora.bounds(tri))
    d.destroy(); out.destroy()


@pytest.mark.parametrize("div_mode", [0, 1])
@pytest.mark.parametrize("host_bounds", [True, False])
def test_aabb_morton(ctx, lb, ora, div_mode, host_bounds):
    for name, tri in mesh_cases(lb).items():
        say(f"aabb_morton {name} div={div_mode} host_bounds={host_bounds}")
        n = len(tri)
        b = lb.wl_bounds(tri)
        ctx.set_div_mode(div_mode)
        d = ctx.to_device(tri)
        aabb = ctx.alloc(32 * n)
        keys = ctx.alloc(4 * n)
        ctx.aabb_morton(d, n, b if host_bounds else None, aabb, keys)
        ga = aabb.read(np.float32, 8 * n).reshape(n, 8)
        gk = keys.read(np.uint32, n)
        wa, wk = ora.aabb_z_idx(tri, b, div_mode)
        assert np.array_equal(gk, wk), name
        assert np.array_equal(ga[:, [0, 1, 2, 4, 5, 6]], wa[:, [0, 1, 2, 4, 5, 6]]), name
        for x in (d, aabb, keys):
            x.destroy()
    ctx.set_div_mode(0)


def test_aabb_morton_caller_bounds_narrower_than_mesh(ctx, lb, ora):
    """OBJ-style input: bounds exclude two huge floor triangles
    (parse-obj.js:65-78), whose quantised coordinates saturate."""
    tri = lb.wl_uniform(31, 5000, 0.01)
    b = lb.wl_bounds(tri)
    floor = np.array([[-1e4, -1e4, 0.01, 3.1415, 1e4, -1e4, 0.01, 3.1415, -1e4, 1e4, 0.01, 3.1415],
                      [1e4, -1e4, 0.01, 3.1415, 1e4, 1e4, 0.01, 3.1415, -1e4, 1e4, 0.01, 3.1415]],
                     np.float32)
    tri = np.concatenate([tri, floor])
    n = len(tri)
    for mode in (0, 1):
        ctx.set_div_mode(mode)
        bvh, otri, sk, pm = ctx.build_arrays(tri, b, keep_topology=True)
        wb, wo, wsk, wpm = ora.build(tri, b, mode)
        assert np.array_equal(sk, wsk) and np.array_equal(pm, wpm)
        assert np.array_equal(bvh, wb)
    ctx.set_div_mode(0)


@pytest.mark.parametrize("bounds_from", ["gpu", "caller"])
def test_planar_mesh_flat_axis(ctx, lb, ora, bounds_from):
    """A planar mesh (every z equal): the flat axis has min == max, so its
    quantised coordinate is 0/0 -> 0 for every triangle, in both division
    modes, from GPU bounds and from caller bounds; a caller box flat on z
    below the mesh saturates the other way."""
    n = 30001
    tri = lb.wl_uniform(57, n, 0.01)
    tri[:, [2, 6, 10]] = np.float32(0.25)
    say(f"planar mesh n={n} bounds={bounds_from}")
    boxes = [None]
    if bounds_from == "caller":
        below = lb.wl_bounds(tri)
        below[2] = below[5] = np.float32(0.125)   # flat box under the plane
        boxes = [lb.wl_bounds(tri), below]
    for b in boxes:
        for mode in (0, 1):
            ctx.set_div_mode(mode)
            bvh, otri, sk, pm = ctx.build_arrays(tri, b, keep_topology=True)
            wb, wo, wsk, wpm = ora.build(tri, b, mode)
            assert np.array_equal(sk, wsk) and np.array_equal(pm, wpm)
            assert np.array_equal(bvh, wb)
            assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32))
    ctx.set_div_mode(0)


def test_non_finite_caller_bounds_are_refused(ctx, lb):
    """NaN or infinite MODEL_BOUNDS are refused before any launch, by the
    build and by the per-stage Morton entry point."""
    tri = lb.wl_uniform(58, 1000, 0.01)
    d = ctx.to_device(tri)
    bvh, otri = ctx.alloc(64 * 1000), ctx.alloc(48 * 1000)
    aabb, keys = ctx.alloc(32 * 1000), ctx.alloc(4 * 1000)
    try:
        for bad in (np.nan, np.inf, -np.inf):
            b = lb.wl_bounds(tri)
            b[4] = bad
            with pytest.raises(lb.LBVHError) as err:
                ctx.build(d, 1000, b, bvh, otri)
            assert err.value.code == lb.ERR_INVALID_ARGUMENT
            with pytest.raises(lb.LBVHError) as err:
                ctx.aabb_morton(d, 1000, b, aabb, keys)
            assert err.value.code == lb.ERR_INVALID_ARGUMENT
        ctx.synchronize()
    finally:
        for x in (d, bvh, otri, aabb, keys):
            x.destroy()


class sort_bits:
    """Fixes the radix-sort digit width of ctx inside a with-block (the
    size-based default after)."""
    def __init__(self, ctx, lb, bits):
        self.ctx, self.lb, self.bits = ctx, lb, bits

    def __enter__(self):
        self.ctx.set_tuning(self.lb.TUNE_SORT_BITS, self.bits)

    def __exit__(self, *exc):
        self.ctx.set_tuning(self.lb.TUNE_SORT_BITS, 0)


@pytest.mark.parametrize("key_bits", [30, 32])
@pytest.mark.parametrize("bits", [8, 10])
@pytest.mark.parametrize("n", [1, 2, 3, 64, 1023, 1024, 1025, 4095, 4096, 4097, 16384,
                               16385, 65535, 65536, 65537, 1000003, 1048576, 1048577,
                               4194303, 4194304, 4194305, 4198400, 4202497])
def test_sort_is_stable(ctx, lb, n, bits, key_bits):
    """Morton-range keys and full u32 keys: the stand-alone sort orders all
    32 bits, as the reference's 16 passes of 2 bits do."""
    say(f"sort n={n} bits={bits} key_bits={key_bits}")
    rng = np.random.default_rng(n + key_bits)
    keys = sort_keys(rng, n, bits=key_bits)
    dk = ctx.to_device(keys)
    di = ctx.alloc(4 * n)
    with sort_bits(ctx, lb, bits):
        ctx.sort(dk, di, n)
    gk = dk.read(np.uint32, n)
    gi = di.read(np.int32, n)
    want = np.argsort(keys, kind="stable")
    assert np.array_equal(gi, want)
    assert np.array_equal(gk, keys[want])
    dk.destroy(); di.destroy()


@pytest.mark.parametrize("offset", [4, 8, 12])
def test_sort_at_unaligned_addresses(ctx, lb, offset):
    """Keys and payloads that do not start on a 16-byte boundary take the
    sort's 4-byte loads; the result is the same stable sort."""
    n = 70000
    say(f"sort at +{offset} bytes")
    rng = np.random.default_rng(offset)
    keys = sort_keys(rng, n)
    pad = offset // 4
    dk = ctx.to_device(np.concatenate([np.zeros(pad, np.uint32), keys]))
    di = ctx.alloc(4 * (n + pad))
    for bits in (8, 10):
        ctx.upload(dk, np.concatenate([np.zeros(pad, np.uint32), keys]))
        with sort_bits(ctx, lb, bits):
            ctx.sort(dk.ptr + offset, di.ptr + offset, n)
        want = np.argsort(keys, kind="stable")
        assert np.array_equal(di.read(np.int32, n + pad)[pad:], want), bits
        assert np.array_equal(dk.read(np.uint32, n + pad)[pad:], keys[want]), bits
    dk.destroy(); di.destroy()


@pytest.mark.parametrize("bits", [8, 10])
def test_sort_heavy_duplicates_and_constant_digits(ctx, lb, bits):
    for n, keys in [(300000, np.zeros(300000, np.uint32)),
                    (300000, np.full(300000, 0x3FFFFFFF, np.uint32)),
                    (123457, (np.arange(123457, dtype=np.uint32) % 3) << 24),
                    (123457, (np.arange(123457, dtype=np.uint32) % 5) << 20),
                    (200000, np.arange(200000, dtype=np.uint32)[::-1].copy()),
                    (300000, np.full(300000, 0xFFFFFFFF, np.uint32)),
                    (123457, (np.arange(123457, dtype=np.uint32) % 4) << 30),
                    (123457, ((np.arange(123457, dtype=np.uint32) % 7) << 29)
                     | (np.arange(123457, dtype=np.uint32)[::-1] & 0xFF)),
                    (4194309, np.zeros(4194309, np.uint32)),
                    (4202497, (np.arange(4202497, dtype=np.uint32) % 255) << 8)]:
        say(f"sort special n={n} bits={bits}")
        dk = ctx.to_device(keys)
        di = ctx.alloc(4 * n)
        with sort_bits(ctx, lb, bits):
            ctx.sort(dk, di, n)
        want = np.argsort(keys, kind="stable")
        assert np.array_equal(di.read(np.int32, n), want)
        assert np.array_equal(dk.read(np.uint32, n), keys[want])
        dk.destroy(); di.destroy()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 100, 4097, 250000])
def test_radix_tree(ctx, ora, n):
    say(f"radix_tree n={n}")
    rng = np.random.default_rng(n + 5)
    keys = np.sort(sort_keys(rng, n, bits=int(rng.integers(2, 31))))
    dk = ctx.to_device(keys)
    dp = ctx.alloc(8 * n)
    ctx.radix_tree(dk, n, dp)
    assert np.array_equal(dp.read(np.int32, 2 * n), ora.radix_tree(keys))
    dk.destroy(); dp.destroy()


def test_stage_pipeline_matches_oracle(ctx, lb, ora):
    """The reference's stage sequence (build-bvh.js:24-56) through the
    per-stage entry points."""
    for name, tri in mesh_cases(lb).items():
        say(f"stage pipeline {name}")
        n = len(tri)
        b = lb.wl_bounds(tri)
        d = ctx.to_device(tri)
        aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
        idx, par = ctx.alloc(4 * n), ctx.alloc(8 * n)
        bvh, otri = ctx.alloc(64 * n), ctx.alloc(48 * n)
        ctx.aabb_morton(d, n, b, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.rearrange(d, idx, n, otri)
        wb, wo, wsk, wpm = ora.build(tri, b, 0)
        assert np.array_equal(keys.read(np.uint32, n), wsk), name
        assert np.array_equal(idx.read(np.int32, n), wpm), name
        assert np.array_equal(par.read(np.int32, 2 * n), ora.radix_tree(wsk)), name
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb), name
        assert np.array_equal(otri.read(np.uint32, 12 * n), wo.view(np.uint32).ravel()), name
        for x in (d, aabb, keys, idx, par, bvh, otri):
            x.destroy()


@pytest.mark.parametrize("n", [1024, 1025, 32768, 32769, 65537, 600001, 1048577])
def test_up_pass_at_level_boundaries(ctx, lb, ora, n):
    """The per-stage up pass leaves each 1,024-leaf block's unfinished
    subtrees as items and merges the items of 32 blocks per launch (in
    rounds of at most 1,024 items) until one block holds the root: 0, 1, 2
    and 3 item levels and ragged last groups."""
    say(f"up pass levels n={n}")
    tri = lb.wl_clustered(n, n, 0.05) if n % 2 else lb.wl_uniform(n, n, 0.01)
    b = lb.wl_bounds(tri)
    d = ctx.to_device(tri)
    aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
    idx, par, bvh = ctx.alloc(4 * n), ctx.alloc(8 * n), ctx.alloc(64 * n)
    try:
        ctx.aabb_morton(d, n, b, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.synchronize()
        wb, _, _, _ = ora.build(tri, b, 0)
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb)
    finally:
        for x in (d, aabb, keys, idx, par, bvh):
            x.destroy()


def test_up_pass_all_equal_keys(ctx, lb, ora):
    """20,000 copies of one triangle: every key equal, the tree splits by
    position only."""
    say("up pass all-equal n=20000")
    tri = np.tile(lb.wl_uniform(31, 1, 0.01), (20000, 1))
    n = len(tri)
    b = lb.wl_bounds(tri)
    d = ctx.to_device(tri)
    aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
    idx, par, bvh = ctx.alloc(4 * n), ctx.alloc(8 * n), ctx.alloc(64 * n)
    try:
        ctx.aabb_morton(d, n, b, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.synchronize()
        wb, _, _, _ = ora.build(tri, b, 0)
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb)
    finally:
        for x in (d, aabb, keys, idx, par, bvh):
            x.destroy()


@pytest.mark.parametrize("kind", ["zeros", "minus_one", "self_loops", "index_out_of_range",
                                  "one_block"])
def test_up_pass_reports_a_parents_array_that_is_not_a_tree(ctx, lb, ora, kind):
    """Inputs that are not a radix tree over the leaves end every climb in
    bounds and flag the context, and lbvh_synchronize reports a device error;
    the flag is cleared and the next valid call succeeds.
      zeros: every leaf a child of the root (frontiers overflow, third children)
      minus_one: every parent out of range (no climb starts, the root never merges)
      self_loops: each internal node its own parent (third arrivals / step cap)
      index_out_of_range: one permutation entry >= n (the load is clamped)
      one_block: 1,000 leaves with all-(-1) parents, the single-block case"""
    n = 1000 if kind == "one_block" else 5000
    say(f"up pass with a malformed input: {kind} n={n}")
    tri = lb.wl_uniform(41, n, 0.01)
    b = lb.wl_bounds(tri)
    d = ctx.to_device(tri)
    aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
    idx, bvh = ctx.alloc(4 * n), ctx.alloc(64 * n)
    par = ctx.alloc(8 * n)
    try:
        ctx.aabb_morton(d, n, b, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.synchronize()
        good_par = par.read(np.int32, 2 * n)
        good_idx = idx.read(np.int32, n)
        bad_par, bad_idx = good_par.copy(), good_idx
        if kind == "zeros":
            bad_par[:] = 0
        elif kind in ("minus_one", "one_block"):
            bad_par[:] = -1
        elif kind == "self_loops":
            bad_par[1:n - 1] = np.arange(1, n - 1, dtype=np.int32)
        else:
            bad_idx = good_idx.copy()
            bad_idx[n // 2] = n + 7
        par.write(bad_par)
        idx.write(bad_idx)
        ctx.up_pass(idx, aabb, par, n, bvh)
        with pytest.raises(lb.LBVHError) as err:
            ctx.synchronize()
        assert err.value.code == lb.ERR_DEVICE
        assert "not a radix tree" in str(err.value) or "more subtrees" in str(err.value)
        par.write(good_par)
        idx.write(good_idx)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.synchronize()
        wb, _, _, _ = ora.build(tri, b, 0)
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb)
    finally:
        for x in (d, aabb, keys, idx, par, bvh):
            x.destroy()


def test_stage_pipeline_heightfield_16m(ctx, lb, ora):
    """The per-stage sequence over the 16,777,216-triangle heightfield: the up
    pass meets most parents on chip and the ones on paths across its 1,024-leaf
    blocks in HBM; the result is the oracle's."""
    say("stage pipeline heightfield-16m")
    tri = lb.workload("heightfield-16m")
    n = len(tri)
    b = lb.wl_bounds(tri)
    d = ctx.to_device(tri)
    aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
    idx, par = ctx.alloc(4 * n), ctx.alloc(8 * n)
    bvh, otri = ctx.alloc(64 * n), ctx.alloc(48 * n)
    try:
        ctx.aabb_morton(d, n, b, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.rearrange(d, idx, n, otri)
        say("oracle heightfield-16m with caller bounds")
        wb, wo, wsk, wpm = ora.build(tri, b, 0)
        assert np.array_equal(keys.read(np.uint32, n), wsk)
        assert np.array_equal(idx.read(np.int32, n), wpm)
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb)
        assert np.array_equal(otri.read(np.uint32, 12 * n), wo.view(np.uint32).ravel())
    finally:
        for x in (d, aabb, keys, idx, par, bvh, otri):
            x.destroy()


# ---------------------------------------------------------------- fused build

@pytest.mark.parametrize("div_mode", [0, 1])
def test_build_matches_oracle(ctx, lb, ora, div_mode):
    ctx.set_div_mode(div_mode)
    for name, tri in mesh_cases(lb).items():
        for with_bounds in (True, False):
            say(f"build {name} div={div_mode} bounds={with_bounds}")
            b = lb.wl_bounds(tri) if with_bounds else None
            bvh, otri, sk, pm = ctx.build_arrays(tri, b, keep_topology=True)
            wb, wo, wsk, wpm = ora.build(tri, b, div_mode)
            assert np.array_equal(sk, wsk), name
            assert np.array_equal(pm, wpm), name
            assert np.array_equal(bvh, wb), name
            assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32)), name
    ctx.set_div_mode(0)


@pytest.mark.parametrize("div_mode", [0, 1])
def test_build_clustered_1m(ctx, lb, ora, div_mode):
    say(f"build clustered-1m div={div_mode}")
    ctx.set_div_mode(div_mode)
    tri = lb.workload("clustered-1m")
    b = lb.wl_bounds(tri)
    bvh, otri, sk, pm = ctx.build_arrays(tri, b, keep_topology=True)
    wb, wo, wsk, wpm = ora.build(tri, b, div_mode)
    ctx.set_div_mode(0)
    assert np.array_equal(sk, wsk)
    assert np.array_equal(pm, wpm)
    assert np.array_equal(bvh, wb)
    assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32))


@pytest.mark.parametrize("tile", [256, 512, 1024])
def test_build_independent_of_tree_tile(ctx, lb, ora, tile):
    """The fused up pass's tile size changes which splits merge in LDS and
    which through HBM, never the result."""
    say(f"tree tile {tile}")
    ctx.set_tuning(lb.TUNE_TREE_TILE, tile)
    try:
        for name in ("clustered100k", "allsame777", "n257", "n2"):
            tri = mesh_cases(lb)[name]
            bvh = ctx.build_arrays(tri, None)[0]
            assert np.array_equal(bvh, ora.build(tri, None, 0)[0]), name
        tri = lb.workload("clustered-1m")
        assert np.array_equal(ctx.build_arrays(tri, None)[0], ora.build(tri, None, 0)[0])
    finally:
        ctx.set_tuning(lb.TUNE_TREE_TILE, 0)


def test_autotune_table_never_changes_results(lb, ora):
    with lb.Context(0) as c:
        times = c.autotune(1 << 10, 1 << 16, reps=2)
        assert sorted(times) == [1 << k for k in range(10, 17)]
        for bucket, per_variant in times.items():
            assert set(per_variant) == {(t, b) for t in (256, 512, 1024) for b in (8, 10)}
            assert all(t > 0 for t in per_variant.values())
            # the fastest variant if it beats the untuned one (tile 1024, the
            # size rule's digit width) by more than 2 %, else the untuned one
            wide = bucket <= 1 << 17 or (1 << 20) < bucket <= 1 << 21
            untuned = (1024, 10 if wide else 8)
            best = min(per_variant, key=per_variant.get)
            if not per_variant[best] < per_variant[untuned] * np.float32(0.98):
                best = untuned
            assert c.get_tuning_for(lb.TUNE_TREE_TILE, bucket) == best[0]
            assert c.get_tuning_for(lb.TUNE_SORT_BITS, bucket) == best[1]
        for n in (1, 700, 5000, 70000):
            tri = lb.wl_clustered(90 + n, n, 0.05)
            assert np.array_equal(c.build_arrays(tri, None)[0], ora.build(tri, None, 0)[0])


@pytest.mark.parametrize("keep_topology", [False, True])
def test_build_wide_digits_match_oracle(ctx, lb, ora, keep_topology):
    """10-bit digits (3 sort passes): the same bytes.  Without caller key and
    permutation buffers the odd pass count ends in the context's other key
    buffer, with them in the caller's."""
    with sort_bits(ctx, lb, 10):
        for name, tri in list(mesh_cases(lb).items()) + [("clustered-1m", lb.workload("clustered-1m"))]:
            say(f"build {name} bits=10 keep_topology={keep_topology}")
            got = ctx.build_arrays(tri, None, keep_topology=keep_topology)
            want = ora.build(tri, None, 0)
            assert np.array_equal(got[0], want[0]), name
            assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32)), name
            if keep_topology:
                assert np.array_equal(got[2], want[2]), name
                assert np.array_equal(got[3], want[3]), name


@pytest.mark.parametrize("n", [16 * 1024, 16 * 1024 + 1, 32 * 1024, 32 * 1024 + 1,
                               64 * 1024, 64 * 1024 + 1, 1024 * 1024, 1024 * 1024 + 1,
                               2048 * 1024, 2048 * 1024 + 1])
def test_build_at_plan_boundaries(ctx, lb, ora, n):
    """Sizes where the plan changes shape (1,024-leaf tiles): 32, 33 and 64
    leaf tiles take one item level (a single block over all of them), 65 take
    two (3 blocks of 32 sources, then one); 2,048 leaf tiles (64, then one)
    and 2,049 (65, then 3, then one).  The sort: up to 2^20 keys in tiles of
    1,024 keys, 16 of them (2^14 keys; the scatter sums the counts itself)
    and 17 (a column-scan kernel); 2^20 + 1 keys in tiles of 4,096.  Both
    digit widths."""
    tri = lb.wl_uniform(n % 9973, n, 0.01)
    want = ora.build(tri, None, 0)
    ctx.set_tuning(lb.TUNE_TREE_TILE, 1024)
    try:
        for bits in (8, 10):
            say(f"plan boundary n={n} bits={bits}")
            with sort_bits(ctx, lb, bits):
                got = ctx.build_arrays(tri, None)
            assert np.array_equal(got[0], want[0]), bits
            assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32)), bits
    finally:
        ctx.set_tuning(lb.TUNE_TREE_TILE, 0)


def test_every_entry_point_with_no_triangles(ctx, lb):
    """N = 0: zero-size work, no dispatch (build-bvh.js with NUM_TRIS = 0).
    Every entry point returns OK and leaves its outputs untouched, except
    the bounds of an empty mesh, which do not exist (invalid argument)."""
    sentinel = np.full(64, 0x5A5A5A5A, np.uint32)
    bufs = [ctx.to_device(sentinel) for _ in range(6)]
    tri, a, b, c, d, e = bufs
    with pytest.raises(lb.LBVHError) as err:
        ctx.bounds(tri, 0, a)
    assert err.value.code == lb.ERR_INVALID_ARGUMENT
    ctx.aabb_morton(tri, 0, None, a, b)
    ctx.aabb_morton(tri, 0, [0, 0, 0, 1, 1, 1], a, b)
    ctx.sort(a, b, 0)
    ctx.radix_tree(a, 0, b)
    ctx.up_pass(a, b, c, 0, d)
    ctx.rearrange(tri, a, 0, b)
    ctx.build(tri, 0, None, a, b, c, d)
    ctx.build(tri, 0, [0, 0, 0, 1, 1, 1], a, b)
    ctx.refit(tri, 0, a, b, c, d)
    ctx.synchronize()
    for x in bufs:
        assert np.array_equal(x.read(np.uint32, 64), sentinel)
        x.destroy()


@pytest.mark.parametrize("n", [1, 2, 1000, 65536, 300001])
@pytest.mark.parametrize("bounds_from", ["gpu", "caller"])
def test_repeated_builds_into_the_same_buffers(ctx, lb, ora, n, bounds_from):
    """The same build (buffers, size, options, bounds) six times in a row,
    the triangles rewritten in place between builds and another build in
    between once: every result is the oracle's, whatever the scratch and
    outputs held from the build before."""
    meshes = [lb.wl_uniform(100 + k, n, 0.01) for k in range(2)]
    b = None
    if bounds_from == "caller":
        lo = np.minimum(lb.wl_bounds(meshes[0]), lb.wl_bounds(meshes[1]))
        hi = np.maximum(lb.wl_bounds(meshes[0]), lb.wl_bounds(meshes[1]))
        b = np.concatenate([lo[:3], hi[3:]]).astype(np.float32)
    want = [ora.build(m, b, 0) for m in meshes]
    dtri = ctx.to_device(meshes[0])
    bvh, otri = ctx.alloc(64 * n), ctx.alloc(48 * n)
    sk, pm = ctx.alloc(4 * n), ctx.alloc(4 * n)
    other = lb.wl_uniform(7, 777, 0.01)
    try:
        for rep in range(6):
            k = rep % 2
            say(f"repeated build n={n} bounds={bounds_from} rep={rep}")
            ctx.upload(dtri, meshes[k])
            ctx.build(dtri, n, b, bvh, otri, sk, pm)
            wb, wo, wsk, wpm = want[k]
            assert np.array_equal(sk.read(np.uint32, n), wsk), rep
            assert np.array_equal(pm.read(np.int32, n), wpm), rep
            assert np.array_equal(bvh.read(np.uint8, 64 * n), wb), rep
            assert np.array_equal(otri.read(np.uint32, 12 * n),
                                  wo.view(np.uint32).reshape(-1)), rep
            if rep == 3:  # a different build between two replays
                assert np.array_equal(ctx.build_arrays(other, None)[0],
                                      ora.build(other, None, 0)[0])
    finally:
        for x in (dtri, bvh, otri, sk, pm):
            x.destroy()


def sweep_mesh(lb, rng, n):
    """A mesh of about n triangles, of a kind drawn from rng: uniform boxes of
    a random size, clustered with a random duplicate share, a small
    heightfield, or a handful of distinct triangles repeated."""
    kind = ("uniform", "clustered", "heightfield", "repeated")[rng.integers(4)]
    seed = int(rng.integers(1 << 30))
    if kind == "uniform":
        tri = lb.wl_uniform(seed, n, float(10 ** rng.uniform(-4, -1)))
    elif kind == "clustered":
        tri = lb.wl_clustered(seed, n, float(rng.uniform(0.0, 0.5)))
    elif kind == "heightfield":
        nx = max(1, int(np.sqrt(n / 2) * rng.uniform(0.5, 2.0)))
        tri = lb.wl_heightfield(seed, nx, max(1, n // (2 * nx)))
    else:
        base = lb.wl_uniform(seed, int(rng.integers(1, 9)), 0.05)
        tri = base[rng.integers(0, base.shape[0], size=n)]
    return kind, np.ascontiguousarray(tri)


@pytest.mark.parametrize("seed", range(32))
def test_build_sweep_matches_oracle(ctx, lb, ora, seed):
    """Seeded draws over what the fused build's code paths depend on: the
    size (log-uniform over 1 .. 2,000,000, so tile, sort-plan and item-level
    boundaries fall anywhere), the mesh kind, the leaf tile, the digit
    width, the division mode, and the bounds (GPU, the mesh's own, a box
    narrower than the mesh so keys saturate, or a wider one).  Keys,
    permutation, nodes and O_TRI must be the oracle's, bit for bit."""
    rng = np.random.default_rng(4000 + seed)
    n = max(1, int(round(np.exp(rng.uniform(0.0, np.log(2000000.0))))))
    kind, tri = sweep_mesh(lb, rng, n)
    n = tri.shape[0]
    tile = int(rng.choice([256, 512, 1024]))
    bits = int(rng.choice([8, 10]))
    div_mode = int(rng.integers(2))
    bounds_from = ("gpu", "own", "narrow", "wide")[rng.integers(4)]
    b = None
    if bounds_from != "gpu":
        own = lb.wl_bounds(tri).astype(np.float64)
        mid, half = (own[:3] + own[3:]) / 2, (own[3:] - own[:3]) / 2
        scale = {"own": None, "narrow": 0.8, "wide": 1.5}[bounds_from]
        b = own if scale is None else np.concatenate([mid - scale * half,
                                                      mid + scale * half])
        b = b.astype(np.float32)
    say(f"sweep seed={seed} n={n} kind={kind} tile={tile} bits={bits} "
        f"div={div_mode} bounds={bounds_from}")
    ctx.set_tuning(lb.TUNE_TREE_TILE, tile)
    ctx.set_tuning(lb.TUNE_SORT_BITS, bits)
    ctx.set_div_mode(div_mode)
    try:
        bvh, otri, sk, pm = ctx.build_arrays(tri, b, keep_topology=True)
    finally:
        ctx.set_tuning(lb.TUNE_TREE_TILE, 0)
        ctx.set_tuning(lb.TUNE_SORT_BITS, 0)
        ctx.set_div_mode(0)
    wb, wo, wsk, wpm = ora.build(tri, b, div_mode)
    assert np.array_equal(sk, wsk)
    assert np.array_equal(pm, wpm)
    assert np.array_equal(bvh, wb)
    assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32))


def test_build_is_deterministic(ctx, lb):
    tri = lb.workload("uniform-64k")
    first = ctx.build_arrays(tri, None)[0]
    for k in range(20):
        say(f"determinism rep {k}")
        assert np.array_equal(ctx.build_arrays(tri, None)[0], first)


_HF16 = {}


@pytest.mark.parametrize("bits", [8, 10])
def test_build_heightfield_16m_full_size(ctx, lb, ora, bits):
    """The 16,777,216-triangle heightfield against the oracle (with either
    sort digit width), plus size-independent checks: sorted keys, a
    permutation, root box = the scene bounds."""
    say(f"build heightfield-16m bits={bits}")
    tri = lb.workload("heightfield-16m")
    n = len(tri)
    with sort_bits(ctx, lb, bits):
        bvh, otri, sk, pm = ctx.build_arrays(tri, None, keep_topology=True)
    assert np.all(sk[1:] >= sk[:-1])
    assert np.array_equal(np.sort(pm), np.arange(n, dtype=np.int32))
    f, i = ora.bvh_fields(bvh, n)
    b = ora.bounds(tri)
    root_min = np.minimum(f[0, 0:3], f[0, 8:11])
    root_max = np.maximum(f[0, 4:7], f[0, 12:15])
    assert np.array_equal(root_min, np.where(b[:3] == 0, np.float32(-1e-8), b[:3]))
    assert np.array_equal(root_max, np.where(b[3:] == 0, np.float32(1e-8), b[3:]))
    say("oracle heightfield-16m")
    if "want" not in _HF16:
        _HF16["want"] = ora.build(tri, None, 0)
    wb, wo, wsk, wpm = _HF16["want"]
    assert np.array_equal(sk, wsk)
    assert np.array_equal(pm, wpm)
    assert np.array_equal(bvh, wb)
    assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32))


def test_build_67m_against_oracle(ctx, lb, ora):
    """2^26 + 3 triangles (3.2 GB in, 4.3 GB of nodes out): four times the
    headline size, where any 32-bit index product (3 float4 per triangle, 4
    int4 per node, 16,385 sort tiles) would wrap if it were computed
    narrower.  Checked against the oracle's multi-threaded build, which the
    CPU tests pin byte for byte to the scalar one."""
    n = (1 << 26) + 3
    say(f"build uniform n={n}")
    tri = lb.wl_uniform(67, n, 0.01)
    bvh, otri, sk, pm = ctx.build_arrays(tri, None, keep_topology=True)
    say("oracle (multi-threaded)")
    wb, wo, wsk, wpm, _ = ora.build_mt(tri, None, 0)
    assert np.array_equal(sk, wsk)
    assert np.array_equal(pm, wpm)
    assert np.array_equal(bvh, wb)
    assert np.array_equal(otri.view(np.uint32), wo.view(np.uint32))


def test_stage_pipeline_67m_against_oracle(ctx, lb, ora):
    """The per-stage sequence at 2^26 + 3 triangles: 65,537 up-pass blocks
    and six item levels above them (8,193, 1,025, 129, 17, 3 and 1 blocks),
    with the GPU bounds reduction, against the oracle's multi-threaded
    build."""
    n = (1 << 26) + 3
    say(f"stage pipeline uniform n={n}")
    tri = lb.wl_uniform(68, n, 0.01)
    d = ctx.to_device(tri)
    aabb, keys = ctx.alloc(32 * n), ctx.alloc(4 * n)
    idx, par, bvh = ctx.alloc(4 * n), ctx.alloc(8 * n), ctx.alloc(64 * n)
    try:
        ctx.aabb_morton(d, n, None, aabb, keys)
        ctx.sort(keys, idx, n)
        ctx.radix_tree(keys, n, par)
        ctx.up_pass(idx, aabb, par, n, bvh)
        ctx.synchronize()
        say("oracle (multi-threaded)")
        wb, _, wsk, wpm, _ = ora.build_mt(tri, None, 0)
        assert np.array_equal(keys.read(np.uint32, n), wsk)
        assert np.array_equal(idx.read(np.int32, n), wpm)
        assert np.array_equal(bvh.read(np.uint8, 64 * n), wb)
    finally:
        for x in (d, aabb, keys, idx, par, bvh):
            x.destroy()


# ---------------------------------------------------------------- refit

def test_refit_matches_oracle(ctx, lb, ora):
    for n in (1, 2, 3, 1000, 65536, 300001):
        say(f"refit n={n}")
        tri = lb.wl_clustered(40 + n, n, 0.05)
        bvh, otri, sk, pm = ctx.build_arrays(tri, None, keep_topology=True)
        moved = lb.wl_jitter(41 + n, tri, 0.001)
        d = ctx.to_device(moved)
        dk, dp = ctx.to_device(sk), ctx.to_device(pm)
        rb, ro = ctx.alloc(64 * n), ctx.alloc(48 * n)
        ctx.refit(d, n, dk, dp, rb, ro)
        wb, wo = ora.refit(moved, sk, pm)
        assert np.array_equal(rb.read(np.uint8, 64 * n), wb)
        assert np.array_equal(ro.read(np.uint32, 12 * n), wo.view(np.uint32).ravel())
        # refit with the original vertices reproduces the build
        d2 = ctx.to_device(tri)
        ctx.refit(d2, n, dk, dp, rb, ro)
        assert np.array_equal(rb.read(np.uint8, 64 * n), bvh)
        for x in (d, dk, dp, rb, ro, d2):
            x.destroy()


# ---------------------------------------------------------------- budget

def test_budget_refusal_leaves_context_usable(lb):
    n = 200000
    tri = lb.wl_uniform(50, n, 0.01)
    user = (48 + 64 + 48) * n
    # room for the inputs and outputs plus 1 MiB: the build's scratch
    # (16 bytes per triangle for keys and indices alone) cannot fit
    with lb.Context(0, hbm_budget=user + (1 << 20) + (64 << 10)) as c:
        d = c.to_device(tri)
        bvh, otri = c.alloc(64 * n), c.alloc(48 * n)
        with pytest.raises(lb.LBVHError) as e:
            c.build(d, n, None, bvh, otri)
        assert e.value.code == 3
        for x in (d, bvh, otri):
            x.destroy()
        small = lb.wl_uniform(51, 1000, 0.01)
        got = c.build_arrays(small, None)[0]
    with lb.Context(0) as c2:
        assert np.array_equal(got, c2.build_arrays(small, None)[0])


# ---------------------------------------------------------------- consumer check

def _rays(seed, count, top=None):
    rng = np.random.default_rng(seed)
    rays = np.zeros((count, 8), np.float32)
    if top is None:
        o = rng.uniform(-0.5, 1.5, (count, 3)).astype(np.float32)
        t = rng.uniform(0, 1, (count, 3)).astype(np.float32)
        rays[:, 0:3], rays[:, 4:7] = o, t - o
    else:  # straight down onto a heightfield, plus some oblique rays
        rays[:, 0] = rng.uniform(0, 1, count)
        rays[:, 1] = rng.uniform(0, top, count)
        rays[:, 2] = 1.0
        rays[:, 4:7] = [0, 0, -1]
        k = count // 2
        rays[:k, 4:7] = rng.uniform(-0.3, 0.3, (k, 3)).astype(np.float32)
        rays[:k, 6] = -1
    return rays


@pytest.mark.parametrize("n", [1, 2, 3, 5000, 100003])
def test_raycast_over_gpu_bvh(ctx, lb, ora, n):
    """The path tracer's traversal (pathtracer.js:495-584) over the GPU-built
    BVH and O_TRI finds the same closest hits as testing every triangle."""
    say(f"raycast n={n}")
    tri = lb.wl_clustered(80 + n, n, 0.05)
    bvh, otri = ctx.build_arrays(tri, None)
    rays = _rays(n, 4096)
    got = ctx.raycast_arrays(bvh, otri, rays)
    want = ora.raycast(bvh, otri, rays)
    bad = np.where((got != want).any(axis=1))[0]
    for q in bad[:8]:
        say(f"ray {q}: {rays[q].tolist()} gpu {got[q].tolist()} cpu {want[q].tolist()}")
    assert len(bad) == 0, f"{len(bad)} of {len(rays)} rays differ"
    if n <= 5000:
        assert np.array_equal(got, ora.raycast_bruteforce(otri, rays))
    else:
        sample = rays[:256]
        assert np.array_equal(got[:256], ora.raycast_bruteforce(otri, sample))


def test_raycast_heightfield_16m(ctx, lb, ora):
    say("raycast heightfield-16m")
    tri = lb.workload("heightfield-16m")
    bvh, otri = ctx.build_arrays(tri, None)
    rays = _rays(5, 8192, top=0.5)
    got = ctx.raycast_arrays(bvh, otri, rays)
    # every vertical ray from above the domain hits the surface (oblique
    # ones may leave the 1 x 0.5 domain)
    assert np.all(got[4096:, 3] > 0)
    want = ora.raycast(bvh, otri, rays)
    bad = np.where((got != want).any(axis=1))[0]
    for q in bad[:8]:
        say(f"ray {q}: {rays[q].tolist()} gpu {got[q].tolist()} cpu {want[q].tolist()}")
    assert len(bad) == 0, f"{len(bad)} of {len(rays)} rays differ"
