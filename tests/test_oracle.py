"""Pins the CPU oracle (oracle/lbvh_oracle.c) before it is trusted.

The reference ships no tests, golden vectors or fixtures, so each oracle
function is checked against what the mathematics fixes: known answers read
off the reference's constants, independent numpy restatements, numpy's
stable sort, and brute-force tree / box constructions that share no code with
the restatement.  Each check is chosen so that a plausible slip (a swapped
axis, a dropped tie-break, an off-by-one in the split search, a missing
perturbation) fails it.
"""
import numpy as np
import pytest


def interleave(x, y, z):
    """Morton code by explicit bit placement: x_k -> 3k, y_k -> 3k+1,
    z_k -> 3k+2 (aabb-z-idx.js:197-199)."""
    out = 0
    for k in range(10):
        out |= ((x >> k) & 1) << (3 * k)
        out |= ((y >> k) & 1) << (3 * k + 1)
        out |= ((z >> k) & 1) << (3 * k + 2)
    return out


def np_keys(tri, b, div_mode):
    """Independent numpy float32 restatement of aabb-z-idx.js:186-194."""
    t = np.asarray(tri, np.float32).reshape(-1, 3, 4)[:, :, :3]
    mn = t.min(axis=1)
    mx = t.max(axis=1)
    b = np.asarray(b, np.float32)
    cen = (mx + mn) * np.float32(0.5)
    d = cen - b[:3]
    ext = b[3:] - b[:3]
    if div_mode == 1:
        rel = d / ext
    else:
        rel = d * (np.float32(1.0) / ext)
    v = (rel * np.float32(1023.99)).astype(np.float64)
    q = np.where(v >= 2.0 ** 32, 2.0 ** 32 - 1, np.where(v > 0, np.trunc(v), 0))
    q = q.astype(np.uint64)
    keys = np.zeros(len(q), np.uint64)
    for axis in range(3):
        # split_3 keeps q's bits 0..9 with bits 24..25 folded onto 8..9
        qa = (q[:, axis] & 0x3FF) | ((q[:, axis] >> 16) & 0x300)
        for k in range(10):
            keys |= ((qa >> k) & 1) << (3 * k + axis)
    return mn, mx, keys.astype(np.uint32)


# ---------------------------------------------------------------- Morton

def test_split3_known_answers(ora):
    assert ora.split_3(0) == 0
    assert ora.split_3(1) == 1
    assert ora.split_3(2) == 8
    assert ora.split_3(3) == 9
    assert ora.split_3(1023) == 0x09249249
    assert ora.morton_code(1, 0, 0) == 1
    assert ora.morton_code(0, 1, 0) == 2
    assert ora.morton_code(0, 0, 1) == 4
    assert ora.morton_code(1023, 1023, 1023) == 0x3FFFFFFF
    assert ora.morton_code(512, 0, 0) == 1 << 27


def test_morton_matches_bit_interleave_exhaustive(ora):
    for x in range(1024):
        assert ora.split_3(x) == interleave(x, 0, 0)
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert ora.morton_code(int(x), int(y), int(z)) == interleave(int(x), int(y), int(z))


def test_split3_out_of_range_fold(ora):
    """Quantised values above 1023 (caller bounds narrower than the mesh):
    split_3's masks keep bits 0..9 and fold bits 24..25 onto 8..9."""
    rng = np.random.default_rng(1)
    vals = list(rng.integers(1024, 2 ** 32, size=3000)) + [0xFFFFFFFF, 1 << 24, 1 << 25, 1 << 26]
    for q in vals:
        q = int(q)
        folded = (q & 0x3FF) | ((q >> 16) & 0x300)
        assert ora.split_3(q) == interleave(folded, 0, 0), hex(q)


def test_f32_to_u32_saturating_truncation(ora):
    cases = {0.0: 0, -0.0: 0, -1.0: 0, 0.999: 0, 1.0: 1, 1023.99: 1023,
             1048575.9: 1048575, 4294967040.0: 4294967040, 2.0 ** 32: 0xFFFFFFFF,
             1e30: 0xFFFFFFFF, float("inf"): 0xFFFFFFFF, float("-inf"): 0,
             float("nan"): 0}
    for v, want in cases.items():
        assert ora.f32_to_u32(v) == want, v


@pytest.mark.parametrize("div_mode", [0, 1])
def test_aabb_and_keys_match_numpy(ora, lb, div_mode):
    for tri in (lb.wl_uniform(11, 20000, 0.01), lb.wl_clustered(12, 20000, 0.05)):
        b = ora.bounds(tri)
        aabb, keys = ora.aabb_z_idx(tri, b, div_mode)
        mn, mx, want = np_keys(tri, b, div_mode)
        assert np.array_equal(aabb[:, :3], mn)
        assert np.array_equal(aabb[:, 4:7], mx)
        assert np.array_equal(keys, want)


def test_keys_out_of_range_bounds(ora, lb):
    """The OBJ loader's bounds exclude its two 10000-unit floor triangles
    (parse-obj.js:65-78): their keys come from saturated quantisation."""
    tri = lb.wl_uniform(13, 500, 0.01)
    b = ora.bounds(tri)
    floor = np.array([[-1e4, -1e4, 0.01, 3.1415, 1e4, -1e4, 0.01, 3.1415, -1e4, 1e4, 0.01, 3.1415],
                      [1e4, -1e4, 0.01, 3.1415, 1e4, 1e4, 0.01, 3.1415, -1e4, 1e4, 0.01, 3.1415]],
                     np.float32)
    tri = np.concatenate([tri, floor])
    for mode in (0, 1):
        _, keys = ora.aabb_z_idx(tri, b, mode)
        _, _, want = np_keys(tri, b, mode)
        assert np.array_equal(keys, want)


def test_division_modes_differ_rarely(ora, lb):
    """The two fdiv lowerings disagree only at quantisation boundaries."""
    tri = lb.wl_uniform(lb.PARITY_SEED, 65536, 0.01)
    b = lb.wl_bounds(tri)
    _, k0 = ora.aabb_z_idx(tri, b, 0)
    _, k1 = ora.aabb_z_idx(tri, b, 1)
    diff = int(np.count_nonzero(k0 != k1))
    assert 0 <= diff < 64


def test_bounds_match_f64_caller_bounds(ora, lb):
    tri = lb.wl_clustered(14, 30000, 0.05)
    assert np.array_equal(ora.bounds(tri), lb.wl_bounds(tri))
    v = tri.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 3)
    assert np.array_equal(ora.bounds(tri), np.concatenate([v.min(0), v.max(0)]))


# ---------------------------------------------------------------- sort

@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 257, 4096, 100003])
def test_radix_sort_is_stable_sort(ora, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 64, size=n, dtype=np.uint32) << rng.integers(0, 25, size=n, dtype=np.uint32)
    sk, idx = ora.radix_sort(keys)
    want = np.argsort(keys, kind="stable")
    assert np.array_equal(idx, want)
    assert np.array_equal(sk, keys[want])


# ---------------------------------------------------------------- tree

def tree_invariants(par, n):
    if n < 2:
        assert np.all(par == 0)
        return
    internal = par[:n - 1]
    leaves = par[n:]
    assert par[0] == 0 and par[n - 1] == 0
    assert np.all((internal >= 0) & (internal < n - 1))
    assert np.all((leaves >= 0) & (leaves < n - 1))
    # every internal node has exactly two children
    counts = np.bincount(np.concatenate([internal[1:], leaves]), minlength=n - 1)
    assert np.all(counts == 2)


def test_tree_known_answers(ora):
    k = np.array([0, 1, 2, 3], np.uint32)
    assert list(ora.radix_tree(k)) == [0, 0, 0, 0, 1, 1, 2, 2]
    # all-equal keys: the index tie-break gives the same shape
    k = np.array([7, 7, 7, 7], np.uint32)
    assert list(ora.radix_tree(k)) == [0, 0, 0, 0, 1, 1, 2, 2]
    k = np.array([5, 5, 5], np.uint32)
    assert list(ora.radix_tree(k)) == [0, 0, 0, 1, 1, 0]
    assert list(ora.radix_tree(np.array([9, 9], np.uint32))) == [0, 0, 0, 0]
    assert list(ora.radix_tree(np.array([9], np.uint32))) == [0, 0]


def test_dif_loop_equals_clz(ora):
    rng = np.random.default_rng(2)
    for a, b in rng.integers(0, 2 ** 32, size=(3000, 2), dtype=np.uint64):
        a, b = int(a), int(b)
        for x, y in ((a, b), (a, a), (a, a ^ 1), (a, a ^ (1 << 31))):
            assert ora.dif(x, y, True) == ora.dif(x, y, False)


@pytest.mark.parametrize("seed", range(12))
def test_tree_faithful_textbook_bruteforce_agree(ora, seed):
    rng = np.random.default_rng(100 + seed)
    for n in [1, 2, 3, 4, 5, 8, 17, 64, 100, 257, 1000]:
        bits = int(rng.integers(1, 31))
        keys = np.sort(rng.integers(0, 2 ** bits, size=n, dtype=np.uint64)).astype(np.uint32)
        a = ora.radix_tree(keys, faithful=True)
        b = ora.radix_tree(keys, faithful=False)
        c = ora.bruteforce_tree(keys)
        assert np.array_equal(a, b), (n, bits)
        assert np.array_equal(b, c), (n, bits)
        tree_invariants(b, n)


# ---------------------------------------------------------------- up pass

def mesh_with_zeros(rng, n):
    tri = rng.uniform(-1, 1, size=(n, 12)).astype(np.float32)
    tri[:, 3::4] = np.float32(3.1415)
    # exact zeros (and negative zeros) on many coordinates
    mask = rng.random((n, 12)) < 0.2
    mask[:, 3::4] = False
    tri[mask] = np.float32(0.0)
    neg = rng.random((n, 12)) < 0.05
    neg[:, 3::4] = False
    tri[neg] = np.float32(-0.0)
    return tri


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 500, 3000])
def test_up_pass_matches_bruteforce_boxes(ora, lb, n):
    rng = np.random.default_rng(n)
    for tri in (mesh_with_zeros(rng, n), lb.wl_clustered(n, n, 0.3)):
        b = ora.bounds(tri)
        aabb, keys = ora.aabb_z_idx(tri, b, 0)
        sk, idx = ora.radix_sort(keys)
        par = ora.radix_tree(sk)
        got = ora.up_pass(idx, aabb, par, canonical=True)
        want = ora.bruteforce_bvh(tri, idx, sk)
        assert np.array_equal(got, want)


def test_up_pass_reference_order_is_canonical_up_to_swap(ora, lb):
    tri = lb.wl_uniform(5, 2000, 0.01)
    b = ora.bounds(tri)
    aabb, keys = ora.aabb_z_idx(tri, b, 0)
    sk, idx = ora.radix_sort(keys)
    par = ora.radix_tree(sk)
    raw = ora.up_pass(idx, aabb, par, canonical=False)
    f, i = ora.bvh_fields(raw, 2000)
    assert np.all(i[:-1, 7] == 2) and np.all(i[:-1, 15] == 0)
    canon = ora.canonicalize(raw, 2000)
    fc, ic = ora.bvh_fields(canon, 2000)
    num = lambda c: np.where(c >= 0, c, -(c + 1))
    assert np.all(num(ic[:-1, 3]) + 1 == num(ic[:-1, 11]))
    assert np.all(raw[-64:] == 0)


def test_leaf_zero_perturbation(ora):
    """A box component equal to +-0 becomes -1e-8f (min) / +1e-8f (max)
    (bvh-up-pass.js:209-210); nonzero components are untouched."""
    tri = np.array([[0.0, 1.0, -0.0, 3.1415, 0.5, 2.0, -0.0, 3.1415, 0.25, 1.5, -0.0, 3.1415],
                    [-1.0, 0.0, 0.5, 3.1415, -2.0, 0.0, 0.75, 3.1415, -1.5, 0.0, 1.0, 3.1415]],
                   np.float32)
    bvh, _, _, pm = ora.build(tri, [-2, 0, 0, 1, 2, 1])
    f, i = ora.bvh_fields(bvh, 2)
    e = np.float32(1e-8)
    boxes = {int(i[0, 3]): (f[0, 0:3], f[0, 4:7]), int(i[0, 11]): (f[0, 8:11], f[0, 12:15])}
    mn0, mx0 = boxes[-(int(np.where(pm == 0)[0][0]) + 1)]
    assert list(mn0) == [np.float32(0.0) - e, 1.0, -e]
    assert list(mx0) == [0.5, 2.0, e]
    mn1, mx1 = boxes[-(int(np.where(pm == 1)[0][0]) + 1)]
    assert list(mn1) == [-2.0, -e, 0.5]
    assert list(mx1) == [-1.0, e, 1.0]


def test_single_triangle_node(ora):
    tri = np.array([[0.1, 0.2, 0.3, 3.1415, 0.4, 0.5, 0.6, 3.1415, 0.2, 0.1, 0.9, 3.1415]], np.float32)
    bvh, otri, sk, pm = ora.build(tri)
    f, i = ora.bvh_fields(bvh, 1)
    assert list(i[0, [3, 7]]) == [-1, 1]
    assert np.allclose(f[0, 0:3], [0.1, 0.1, 0.3]) and np.allclose(f[0, 4:7], [0.4, 0.5, 0.9])
    assert np.all(i[0, 8:] == 0)
    assert list(pm) == [0]


# ---------------------------------------------------------------- rearrange

def test_rearrange_is_gather(ora, lb):
    tri = lb.wl_uniform(6, 1000, 0.01)
    idx = np.random.default_rng(6).permutation(1000).astype(np.int32)
    assert np.array_equal(ora.rearrange(tri, idx), tri[idx])


def test_refit_with_unchanged_vertices_equals_build(ora, lb):
    tri = lb.wl_clustered(7, 5000, 0.05)
    bvh, otri, sk, pm = ora.build(tri)
    rbvh, rotri = ora.refit(tri, sk, pm)
    assert np.array_equal(rbvh, bvh) and np.array_equal(rotri, otri)


def test_refit_boxes_match_bruteforce(ora, lb):
    tri = lb.wl_clustered(8, 3000, 0.05)
    bvh, otri, sk, pm = ora.build(tri)
    moved = lb.wl_jitter(9, tri, 0.001)
    rbvh, rotri = ora.refit(moved, sk, pm)
    assert np.array_equal(rbvh, ora.bruteforce_bvh(moved, pm, sk))
    assert np.array_equal(rotri, moved[pm])


# ---------------------------------------------------------------- consumer

def make_rays(seed, count, lo=-0.5, hi=1.5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (count, 3)).astype(np.float32)
    t = rng.uniform(0, 1, (count, 3)).astype(np.float32)
    rays = np.zeros((count, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = t - o
    rays[: count // 8, 4:7] = [0, 0, -1]     # axis-aligned: exercises d += 1e-5
    rays[: count // 8, 0:3] = np.c_[rng.uniform(0, 1, (count // 8, 2)), np.full(count // 8, 2)]
    return rays


@pytest.mark.parametrize("n", [1, 2, 3, 300, 20000])
def test_raycast_traversal_equals_bruteforce(ora, lb, n):
    """The path tracer's traversal over the oracle's BVH finds exactly the
    closest hit that testing every triangle finds."""
    tri = lb.wl_clustered(70 + n, n, 0.05)
    bvh, otri, sk, pm = ora.build(tri)
    rays = make_rays(n, 3000)
    h1 = ora.raycast(bvh, otri, rays)
    h2 = ora.raycast_bruteforce(otri, rays)
    assert np.array_equal(h1, h2)
    if n >= 300:
        assert (h1[:, 3] > 0).sum() >= (50 if n >= 20000 else 5)


# ---------------------------------------------------------------- multi-threaded CPU build

@pytest.mark.parametrize("n", [1, 2, 3, 7, 255, 4097, 30000])
def test_build_mt_matches_build(ora, n):
    """The multi-threaded CPU build (the CPU baseline) gives the oracle's
    bytes: sorted keys, permutation, canonical BVH and O_TRI."""
    rng = np.random.default_rng(n)
    c = rng.random((n, 3)).astype(np.float32)
    tri = np.zeros((n, 12), np.float32)
    for v in range(3):
        tri[:, 4 * v:4 * v + 3] = c + rng.uniform(-0.01, 0.01, (n, 3)).astype(np.float32)
        tri[:, 4 * v + 3] = 3.1415
    if n > 100:
        tri[n // 2:n // 2 + 50] = tri[:50]      # duplicate keys
        tri[10, 0:3] = 0.0                      # exact zeros in a leaf box
    for div in (ora.DIV_MUL_RCP, ora.DIV_IEEE):
        want = ora.build(tri, None, div)
        got = ora.build_mt(tri, None, div)
        for w, g in zip(want, got[:4]):
            assert np.array_equal(np.asarray(w).view(np.uint8), np.asarray(g).view(np.uint8))
        assert got[4] >= 1


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 70001])
def test_refit_mt_matches_refit(ora, lb, n):
    """The multi-threaded refit (used to replay the serving loop) gives the
    scalar refit's bytes on jittered vertices, zeros in leaf boxes included."""
    tri = lb.wl_clustered(n + 5, n, 0.05) if n > 2 else lb.wl_uniform(n + 5, n, 0.01)
    _, _, sk, pm = ora.build(tri)
    moved = lb.wl_jitter(n + 6, tri, 0.001)
    if n > 100:
        moved[17, 0:3] = 0.0
    want = ora.refit(moved, sk, pm)
    got = ora.refit_mt(moved, sk, pm)
    assert np.array_equal(want[0], got[0])
    assert np.array_equal(want[1].view(np.uint32), got[1].view(np.uint32))


def _splitmix64_np(x):
    x = x + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def _digest_np(words):
    words = words.reshape(-1)
    with np.errstate(over="ignore"):
        j = np.arange(words.size, dtype=np.uint64)
        w = _splitmix64_np(j) | np.uint64(1)
        return int(np.sum(words.astype(np.uint64) * w, dtype=np.uint64))


def test_digest_matches_numpy_and_sees_single_words(ora, lb):
    """oracle.digest, the serving loop's per-request digest, against an
    independent numpy restatement; changing any one word (or swapping two
    different words) changes it."""
    tri = lb.wl_clustered(3, 5000, 0.05)
    bvh, otri, _, _ = ora.build(tri)
    want = (_digest_np(bvh.view(np.uint32)), _digest_np(otri.view(np.uint32)))
    assert ora.digest(bvh, otri) == want
    # splitmix64 known answer: the first output of the reference generator
    # seeded with 0 (Vigna's splitmix64.c) is 0xE220A8397B1DCDAF
    assert int(_splitmix64_np(np.array([0], np.uint64))[0]) == 0xE220A8397B1DCDAF
    rng = np.random.default_rng(1)
    for _ in range(20):
        b2 = bvh.copy().view(np.uint32)
        j = int(rng.integers(0, b2.size))
        b2[j] ^= np.uint32(1 << int(rng.integers(0, 32)))
        assert ora.digest(b2.view(np.uint8), otri)[0] != want[0]
    o2 = otri.copy().view(np.uint32).reshape(-1)
    o2[[5, 9]] = o2[[9, 5]]
    assert ora.digest(bvh, o2.view(np.float32).reshape(-1, 12))[1] != want[1]
