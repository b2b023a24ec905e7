"""CPU model of the per-stage radix tree as the GPU computes it, checked
against the oracle's restatement of radix-tree.js:137-238.

The Karras tree over sorted keys is the min-Cartesian tree of the gap deltas
d(g) = delta(g, g+1) (radix-tree.js:172-185): the node split at gap g covers
the leaves between the nearest gaps on either side with a smaller delta,
L(g) and R(g) (all nearest smaller values; sentinels -1 and n-1 with delta
-1).  Its parent splits at whichever of the two has the larger delta: at R,
the node is the parent's left child and its Karras id is its right end R; at
L, the right child with id L+1 (the root, both sentinels, is 0).  Its
children are numbered g and g+1 (radix-tree.js:226-235), leaves when their
range is one leaf, so gap g writes exactly the parents slots g / n+g and
g+1 / n+g+1.

The kernel finds L and R in three steps, modelled here with small tiles so
that every path runs: (1) inside a tile, through 64-gap segments; gaps with
no smaller delta inside the tile (the tile's left and right spines) are
deferred; (2) each tile summarises, for every delta value v, its last and
first gap with a smaller delta (read off its spines), and a prefix / suffix
scan over the tiles carries those summaries; (3) the deferred gaps take
their bounds from the scans.
"""
import numpy as np
import pytest


def clz32(x):
    x = x.astype(np.uint64)
    out = np.full(x.shape, 32, np.int64)
    nz = x != 0
    out[nz] = 31 - np.floor(np.log2(x[nz].astype(np.float64))).astype(np.int64)
    return out


def gap_deltas(sk):
    n = len(sk)
    g = np.arange(n - 1, dtype=np.int64)
    a = sk[:-1].astype(np.uint64)
    b = sk[1:].astype(np.uint64)
    x = (a ^ b).astype(np.uint32)
    same = 32 + clz32((g ^ (g + 1)).astype(np.uint32))
    return np.where(x != 0, clz32(x), same)


def ansv_stack(d):
    m = len(d)
    L = np.full(m, -1, np.int64)
    R = np.full(m, m, np.int64)
    st = []
    for i in range(m):
        while st and d[st[-1]] >= d[i]:
            R[st.pop()] = i
        L[i] = st[-1] if st else -1
        st.append(i)
    return L, R


def parents_from_bounds(n, d, L, R):
    par = np.zeros(2 * n, np.int32)
    g = np.arange(n - 1, dtype=np.int64)
    dl = np.where(L >= 0, d[np.maximum(L, 0)], -1)
    dr = np.where(R <= n - 2, d[np.minimum(R, n - 2)], -1)
    ids = np.where(dl > dr, L + 1, R)
    ids[(dl == -1) & (dr == -1)] = 0
    par[np.where(L + 1 == g, n + g, g)] = ids
    par[np.where(g + 1 == R, n + g + 1, g + 1)] = ids
    par[0] = 0
    par[n - 1] = 0
    return par


def tiled_bounds(d, tile, seg):
    """L and R the way the kernel finds them (see the module docstring)."""
    m = len(d)
    tiles = (m + tile - 1) // tile
    L = np.full(m, -2, np.int64)   # -2: deferred to step 3
    R = np.full(m, -2, np.int64)
    last = np.full((tiles, 64), -1, np.int64)       # last gap with d < v
    first = np.full((tiles, 64), m, np.int64)       # first gap with d < v
    for t in range(tiles):
        a, b = t * tile, min(m, (t + 1) * tile)
        segs = [(s, min(b, s + seg)) for s in range(a, b, seg)]
        smin = [d[s:e].min() for s, e in segs]
        for si, (s, e) in enumerate(segs):
            for i in range(s, e):
                v = d[i]
                lt = [j for j in range(s, i) if d[j] < v]
                if lt:
                    L[i] = lt[-1]
                else:
                    for sj in range(si - 1, -1, -1):
                        if smin[sj] < v:
                            s2, e2 = segs[sj]
                            L[i] = [j for j in range(s2, e2) if d[j] < v][-1]
                            break
                gt = [j for j in range(i + 1, e) if d[j] < v]
                if gt:
                    R[i] = gt[0]
                else:
                    for sj in range(si + 1, len(segs)):
                        if smin[sj] < v:
                            s2, e2 = segs[sj]
                            R[i] = [j for j in range(s2, e2) if d[j] < v][0]
                            break
        # summaries from the spines: the right spine (R deferred) holds, for
        # each v, the tile's last gap with d < v: its largest value below v
        rs = [i for i in range(a, b) if R[i] == -2]
        ls = [i for i in range(a, b) if L[i] == -2]
        assert len(rs) <= 64 and len(ls) <= 64
        for v in range(64):
            below = [i for i in rs if d[i] < v]
            if below:
                last[t, v] = max(below, key=lambda i: d[i])
            below = [i for i in ls if d[i] < v]
            if below:
                first[t, v] = max(below, key=lambda i: d[i])
    # scans over the tiles
    P = np.full((tiles, 64), -1, np.int64)
    S = np.full((tiles, 64), m, np.int64)
    for t in range(1, tiles):
        P[t] = np.where(last[t - 1] >= 0, last[t - 1], P[t - 1])
    for t in range(tiles - 2, -1, -1):
        S[t] = np.where(first[t + 1] < m, first[t + 1], S[t + 1])
    for i in range(m):
        t = i // tile
        if L[i] == -2:
            L[i] = P[t, d[i]]
        if R[i] == -2:
            R[i] = S[t, d[i]]
    return L, R


def sorted_keys(seed, n, bits, dup):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 2 ** bits, n, dtype=np.uint64).astype(np.uint32)
    if dup:
        m = rng.random(n) < dup
        k[m] = k[rng.integers(0, n, int(m.sum()))]
    return np.sort(k)


@pytest.mark.parametrize("n,bits,dup", [(2, 30, 0), (3, 30, 0), (5, 30, 0.5), (700, 30, 0.3),
                                        (700, 4, 0), (1000, 0, 0), (3000, 30, 0.05),
                                        (3000, 12, 0.2)])
def test_ansv_parents_equal_oracle(ora, n, bits, dup):
    sk = sorted_keys(n + bits, n, bits, dup)
    want = ora.radix_tree(sk)
    d = gap_deltas(sk)
    L, R = ansv_stack(d)
    assert np.array_equal(parents_from_bounds(n, d, L, R), want)
    Lt, Rt = tiled_bounds(d, tile=64, seg=8)
    assert np.array_equal(Lt, L) and np.array_equal(Rt, R)
