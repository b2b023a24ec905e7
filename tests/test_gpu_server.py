"""Serving loop (lbvh_server_*): results are the same alone and in the mix,
same-mesh requests keep their order, the HBM budget is enforced by LRU
eviction, and failures are reported per request."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def wait_all(srv):
    srv.drain()
    return {c["request_id"]: c for c in srv.poll()}


# (kind, mesh, source, n, seed, param): builds of synthesised uniform meshes
# and jitter refits, several per mesh so ordering matters.
SCRIPT = [
    (0, 1, 2, 70000, 11, 0.01), (0, 2, 2, 3000, 12, 0.01), (0, 3, 2, 250000, 13, 0.01),
    (1, 1, 3, 70000, 21, 0.001), (0, 4, 2, 1, 14, 0.01), (1, 2, 3, 3000, 22, 0.001),
    (1, 1, 3, 70000, 23, 0.001), (0, 5, 2, 9000, 15, 0.01), (1, 3, 3, 250000, 24, 0.001),
    (0, 2, 2, 5000, 16, 0.01), (1, 5, 3, 9000, 25, 0.001), (1, 2, 3, 5000, 26, 0.001),
]


def run_script(lb, ctx, slots, together):
    with lb.Server(0, 0, slots) as srv:
        for (kind, mesh, src, n, seed, param) in SCRIPT:
            srv.submit(kind, mesh, src, n, seed=seed, param=param)
            if not together:
                done = wait_all(srv)
                assert all(c["status"] == 0 for c in done.values())
        if together:
            done = wait_all(srv)
            assert len(done) == len(SCRIPT)
            assert all(c["status"] == 0 for c in done.values()), done
        return {m: srv.read_mesh(ctx, m) for m in {s[1] for s in SCRIPT}}


def test_same_alone_and_in_the_mix(ctx, lb):
    say("server: alone")
    alone = run_script(lb, ctx, 1, together=False)
    say("server: mixed, 4 slots")
    mixed = run_script(lb, ctx, 4, together=True)
    for m in alone:
        assert np.array_equal(alone[m][0], mixed[m][0]), m
        assert np.array_equal(alone[m][1], mixed[m][1]), m


def test_server_results_match_oracle(ctx, lb, ora):
    """A synthesised build equals the host generator's mesh built by the
    oracle; a jitter refit equals the oracle's refit of the jittered
    vertices (read back from the server's O_TRI in leaf order)."""
    with lb.Server(0, 0, 2) as srv:
        srv.submit(lb.REQ_BUILD, 7, lb.SRC_UNIFORM, 40000, seed=99, param=0.01)
        assert all(c["status"] == 0 for c in wait_all(srv).values())
        bvh, otri = srv.read_mesh(ctx, 7)
        tri = lb.wl_uniform(99, 40000, 0.01)
        wb, wo, wsk, wpm = ora.build(tri, None, 0)
        assert np.array_equal(bvh, wb)
        assert np.array_equal(otri, wo)
        srv.submit(lb.REQ_REFIT, 7, lb.SRC_JITTER, 40000, seed=5, param=0.001)
        assert all(c["status"] == 0 for c in wait_all(srv).values())
        rbvh, rotri = srv.read_mesh(ctx, 7)
        moved = np.empty_like(tri)
        moved[wpm] = rotri  # O_TRI[i] = tri[perm[i]]
        assert not np.array_equal(moved, tri)
        wb2, wo2 = ora.refit(moved, wsk, wpm)
        assert np.array_equal(rbvh, wb2)
        assert np.array_equal(rotri, wo2)


def test_budget_eviction_and_failures(lb):
    n = 100000
    per_mesh = 168 * n
    budget = 3 * per_mesh + 40 * n + (4 << 20)
    with lb.Server(0, budget, 2) as srv:
        for m in range(6):
            srv.submit(lb.REQ_BUILD, m, lb.SRC_UNIFORM, n, seed=m, param=0.01)
        done = wait_all(srv)
        assert all(c["status"] == 0 for c in done.values())
        st = srv.stats()
        assert st["evictions"] >= 3 and st["resident_meshes"] <= 3
        assert st["peak_bytes"] <= budget
        # the most recent mesh is resident, the first was evicted
        srv.mesh(5)
        with pytest.raises(lb.LBVHError):
            srv.mesh(0)
        # a request that can never fit, a refit of an evicted mesh, a refit
        # with the wrong size
        big = srv.submit(lb.REQ_BUILD, 100, lb.SRC_UNIFORM, 10 * n, seed=1, param=0.01)
        gone = srv.submit(lb.REQ_REFIT, 0, lb.SRC_JITTER, n, seed=1, param=0.001)
        wrong = srv.submit(lb.REQ_REFIT, 5, lb.SRC_JITTER, n + 1, seed=1, param=0.001)
        ok = srv.submit(lb.REQ_REFIT, 5, lb.SRC_JITTER, n, seed=1, param=0.001)
        done = wait_all(srv)
        assert done[big]["status"] == 3
        assert done[gone]["status"] == 8
        assert done[wrong]["status"] == 2
        assert done[ok]["status"] == 0
        with pytest.raises(lb.LBVHError):
            srv.submit(lb.REQ_BUILD, 1, lb.SRC_UNIFORM, 0, seed=1)
