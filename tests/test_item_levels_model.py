"""The item levels of the tree build (csrc/k_tree.hip, k_items) restated on
the CPU, step for step: the gap deltas as a sparse table of range minima,
each gap's nearest smaller delta on either side by binary lifting, child boxes
from a segment tree over the item boxes, completeness flags and the forest's
roots as the next frontier, with every array access bounds-checked.  A block
takes the frontiers of FAN tiles (all of them, in one last block, once at
most FAN_MAX remain) in rounds of at most T items, each round
starting with the previous round's frontier.  Run level after level over
small meshes (also with rounds far smaller than the kernel's, so that most
blocks take several), it must reproduce the oracle's nodes byte for byte (bvh-up-pass.js:200-280 with the Karras ids of
radix-tree.js:226-235, as the oracle restates them)."""
import numpy as np
import pytest

T = 1024      # items per block (kItemTile)
LOG = 10      # sparse-table levels above 0
FMAX = 128
FAN = 32     # source tiles per block (kFanIn) ...
FAN_MAX = 64  # ... or all of them, up to this many, in one last block (fan_in)
INF = np.float32(np.inf)

def clz(x): return 32 - int(x).bit_length()
def delta_next(a, b, i): return clz(a ^ b) if a != b else 32 + clz(i ^ (i + 1))

class Arr:
    """fixed-size array with bounds checks"""
    def __init__(self, n, fill=0, dtype=np.int64):
        self.a = np.full(n, fill, dtype=dtype)
    def __getitem__(self, i):
        assert 0 <= i < len(self.a), (i, len(self.a)); return self.a[i]
    def __setitem__(self, i, v):
        assert 0 <= i < len(self.a), (i, len(self.a)); self.a[i] = v

def item_level(items, skeys, n, nodes, T=T, LOG=LOG):
    """items: list of (l, r, mn, mx, dl, dr) in leaf order (m <= T).
    Returns the frontier list in the same format."""
    m = len(items)
    assert m <= T
    # P1: loads
    il = Arr(T); ir = Arr(T)
    dm = [Arr(T + 2, 127) for _ in range(LOG + 1)]  # int8 in the kernel; position p = gap + 1
    smn = [Arr(2 * T, INF, np.float32) for _ in range(3)]
    smx = [Arr(2 * T, -INF, np.float32) for _ in range(3)]
    for k in range(m):
        l, r, mn, mx, dl, dr = items[k]
        il[k] = l; ir[k] = r
        for c in range(3):
            smn[c][T + k] = mn[c]; smx[c][T + k] = mx[c]
        dm[0][k + 1] = dr
        if k == 0: dm[0][0] = dl
        else: assert dm[0][k] == dl, "gap delta mismatch between neighbouring items"
    # P2: sparse table over positions [0, m], segment tree over T leaves
    for s in range(1, LOG + 1):
        w = 1 << s
        for p in range(0, m + 2 - w):
            dm[s][p] = min(dm[s - 1][p], dm[s - 1][p + (w >> 1)])
        for i in range(T >> s, T >> (s - 1)):
            for c in range(3):
                smn[c][i] = min(smn[c][2 * i], smn[c][2 * i + 1])
                smx[c][i] = max(smx[c][2 * i], smx[c][2 * i + 1])
    def query(a, b):  # items [a, b]
        assert 0 <= a <= b < m
        qmn = [INF] * 3; qmx = [-INF] * 3
        lo, hi = a + T, b + T + 1
        while lo < hi:
            if lo & 1:
                for c in range(3): qmn[c] = min(qmn[c], smn[c][lo]); qmx[c] = max(qmx[c], smx[c][lo])
                lo += 1
            if hi & 1:
                hi -= 1
                for c in range(3): qmn[c] = min(qmn[c], smn[c][hi]); qmx[c] = max(qmx[c], smx[c][hi])
            lo >>= 1; hi >>= 1
        return np.array(qmn, np.float32), np.array(qmx, np.float32)
    # P3: every internal gap g (0 .. m-2): nearest smaller on each side
    complete = Arr(T, 0); qLa = Arr(T, -1); qRa = Arr(T, -1)
    for g in range(m - 1):
        p = g + 1; d = dm[0][p]
        cur = p
        for s in range(LOG, -1, -1):
            w = 1 << s
            if cur - w >= 0 and dm[s][cur - w] >= d: cur -= w
        qL = cur - 1 if cur > 0 else -1          # position of the nearest smaller
        cur = p + 1
        for s in range(LOG, -1, -1):
            w = 1 << s
            if cur + w - 1 <= m and dm[s][cur] >= d: cur += w
        qR = cur if cur <= m else -1
        if qL >= 0: assert dm[0][qL] < d
        if qR >= 0: assert dm[0][qR] < d
        if qL >= 0 and qR >= 0:
            complete[g] = 1; qLa[g] = qL; qRa[g] = qR
    # the complete nodes, written at their Karras ids
    for g in range(m - 1):
        if not complete[g]: continue
        qL, qR = qLa[g], qRa[g]
        a, b = qL, qR - 1          # first and last item of the node
        lmn, lmx = query(a, g); rmn, rmx = query(g + 1, b)
        Lr, s_, Rr = il[a], ir[g], ir[b]
        dl, dr = dm[0][qL], dm[0][qR]
        root = Lr == 0 and Rr == n - 1
        K = 0 if root else (Rr if dr > dl else Lr)
        lid = -(s_ + 1) if Lr == s_ else s_
        rid = -(s_ + 2) if Rr == s_ + 1 else s_ + 1
        nd = np.zeros(16, np.int32)
        nd[0:3] = lmn.view(np.int32); nd[3] = lid; nd[4:7] = lmx.view(np.int32); nd[7] = 2
        nd[8:11] = rmn.view(np.int32); nd[11] = rid; nd[12:15] = rmx.view(np.int32); nd[15] = 0
        assert (nodes[K] == 0).all(), "node written twice"
        nodes[K] = nd
    # P4: roots of the forest, marked at their first item
    root_of = Arr(T, -1)
    for g in range(m - 1):
        if not complete[g]: continue
        qL, qR = qLa[g], qRa[g]
        a, b = qL, qR - 1
        if il[a] == 0 and ir[b] == n - 1: continue   # the mesh root: no parent
        dl, dr = dm[0][qL], dm[0][qR]
        parent = (qR - 1) if dr > dl else (qL - 1)   # gap index
        if 0 <= parent <= m - 2 and complete[parent]: continue
        assert root_of[a] == -1; root_of[a] = g
    for i in range(m):
        if il[i] == 0 and ir[i] == n - 1: continue   # a whole-mesh item
        dl, dr = dm[0][i], dm[0][i + 1]
        parent = i if dr > dl else i - 1
        if 0 <= parent <= m - 2 and complete[parent]: continue
        assert root_of[i] == -1; root_of[i] = T + i   # an item root
    # P5: frontier in item order
    out = []
    for k in range(m):
        v = root_of[k]
        if v < 0: continue
        if v >= T:
            i = v - T
            mn = np.array([smn[c][T + i] for c in range(3)], np.float32)
            mx = np.array([smx[c][T + i] for c in range(3)], np.float32)
            out.append((il[i], ir[i], mn, mx, dm[0][i], dm[0][i + 1]))
        else:
            g = v; a, b = qLa[g], qRa[g] - 1
            mn, mx = query(a, b)
            out.append((il[a], ir[b], mn, mx, dm[0][a], dm[0][b + 1]))
    assert len(out) <= FMAX, len(out)
    return out

def leaf_level(tri, skeys, perm, n, Tl, nodes):
    """The leaf tiles, by the same method over single-leaf items (the GPU
    leaf kernel climbs instead; both give the same nodes and frontier)."""
    leafbox = []
    for k in range(n):
        t = tri[perm[k]].reshape(3, 4)[:, :3]
        mn = t.min(0).astype(np.float32); mx = t.max(0).astype(np.float32)
        mn = np.where(mn == 0, np.float32(-1e-8), mn); mx = np.where(mx == 0, np.float32(1e-8), mx)
        leafbox.append((mn, mx))
    def dgap_of(r):  # delta at the gap after leaf r
        return delta_next(int(skeys[r]), int(skeys[r + 1]), r) if r + 1 < n else -1
    tiles = -(-n // Tl)
    fronts = []
    for tile in range(tiles):
        ts, te = tile * Tl, min(tile * Tl + Tl, n)
        # leaf tile == item level over single-leaf items with the climb replaced
        # by the same parallel method (equivalent node set); deltas from keys
        items = [(k, k, leafbox[k][0], leafbox[k][1],
                  dgap_of(k - 1) if k > 0 else -1, dgap_of(k)) for k in range(ts, te)]
        fronts.append(item_level(items, skeys, n, nodes))
    return fronts

def item_block(flat, skeys, n, nodes, T_round):
    """One block of an item level: rounds of at most T_round items, the
    previous round's frontier first (k_items)."""
    log = T_round.bit_length() - 1
    assert 1 << log == T_round
    carry, taken = [], 0
    while True:
        fresh = min(len(flat) - taken, T_round - len(carry))
        assert fresh > 0 or taken == len(flat)
        items = carry + flat[taken:taken + fresh]
        front = item_level(items, skeys, n, nodes, T_round, log)
        if taken + fresh == len(flat):
            return front
        taken += fresh
        carry = front

def run(ora, tri, Tl, T_round=T):
    bvh, otri, sk, pm = ora.build(tri)
    n = len(sk)
    nodes = np.zeros((max(n, 1), 16), np.int32)
    fronts = leaf_level(tri, sk, pm, n, Tl, nodes)
    levels = 1
    while len(fronts) > 1 or (fronts and fronts[0]):
        nxt = []
        fan = FAN_MAX if len(fronts) <= FAN_MAX else FAN
        for t in range(0, len(fronts), fan):
            flat = [it for f in fronts[t:t + fan] for it in f]
            for a, b in zip(flat, flat[1:]):
                assert a[1] + 1 == b[0]
            nxt.append(item_block(flat, sk, n, nodes, T_round))
        fronts = nxt; levels += 1
        if len(fronts) == 1 and not fronts[0]: break
        assert levels < 20
    got = nodes.view(np.uint8).ravel()
    ok = np.array_equal(got[:64 * (n - 1)], bvh[:64 * (n - 1)])
    return ok, n, levels



# (name, mesh, items per round)
CASES = [("clustered", lambda lb: lb.wl_clustered(5, 6000, 0.3), T),
         ("uniform", lambda lb: lb.wl_uniform(6, 5000, 0.01), T),
         ("heightfield", lambda lb: lb.wl_heightfield(7, 32, 32), T),
         ("all-equal", lambda lb: np.tile(lb.wl_uniform(9, 1, 0.01), (700, 1)), T),
         ("n2", lambda lb: lb.wl_uniform(3, 2, 0.01), T),
         ("n3", lambda lb: lb.wl_uniform(4, 3, 0.01), T),
         ("multi-level", lambda lb: lb.wl_uniform(11, 33 * 32 * 64 + 5, 0.01), T),
         ("rounds-uniform", lambda lb: lb.wl_uniform(12, 40000, 0.01), 128),
         ("rounds-clustered", lambda lb: lb.wl_clustered(13, 30000, 0.3), 128),
         ("rounds-all-equal", lambda lb: np.tile(lb.wl_uniform(14, 1, 0.01), (9000, 1)), 128)]


@pytest.mark.parametrize("name", [c[0] for c in CASES])
def test_parallel_item_levels_reproduce_oracle(lb, ora, name):
    gen, t_round = {c[0]: c[1:] for c in CASES}[name]
    tri = gen(lb)
    ok, n, levels = run(ora, tri, 64, t_round)
    assert ok, (name, n, levels)
