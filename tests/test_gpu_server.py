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
        # the device's jitter is the host generator's, bit for bit
        assert np.array_equal(moved.view(np.uint32),
                              lb.wl_jitter(5, tri, 0.001).view(np.uint32))
        wb2, wo2 = ora.refit(moved, wsk, wpm)
        assert np.array_equal(rbvh, wb2)
        assert np.array_equal(rotri, wo2)


def test_completion_digests_are_the_oracle_digests(ctx, lb, ora):
    """With digests on, a completion carries the digests of its results as
    the oracle computes them from the bytes read back."""
    with lb.Server(0, 0, 2) as srv:
        srv.set_digest(True)
        a = srv.submit(lb.REQ_BUILD, 3, lb.SRC_UNIFORM, 30001, seed=7, param=0.01)
        srv.drain()
        done = {c["request_id"]: c for c in srv.poll()}
        bvh, otri = srv.read_mesh(ctx, 3)
        assert done[a]["status"] == 0
        assert (done[a]["digest_bvh"], done[a]["digest_tri"]) == ora.digest(bvh, otri)
        b = srv.submit(lb.REQ_REFIT, 3, lb.SRC_JITTER, 30001, seed=8, param=0.001)
        srv.drain()
        done = {c["request_id"]: c for c in srv.poll()}
        bvh2, otri2 = srv.read_mesh(ctx, 3)
        assert not np.array_equal(otri2, otri)
        assert (done[b]["digest_bvh"], done[b]["digest_tri"]) == ora.digest(bvh2, otri2)
        srv.set_digest(False)
        c = srv.submit(lb.REQ_REFIT, 3, lb.SRC_JITTER, 30001, seed=9, param=0.001)
        srv.drain()
        done = {x["request_id"]: x for x in srv.poll()}
        assert done[c]["status"] == 0 and done[c]["digest_bvh"] == 0


@pytest.mark.parametrize("seed", [6, 7, 8, 9, 10, 11])
def test_serving_plan_stays_within_budget(lb, seed):
    """The serving plan of the test below under other seeds, checked for the
    budget only: the device memory the server holds (mesh arena and every
    slot's scratch, counted at each admission) never exceeds it.  A slot's
    scratch may be trimmed while a request is being admitted to it, and a
    build and a refit grow different buffers of one context."""
    requests = 240
    budget = 3 << 30
    plan = lb.wl_serving_plan(np.random.default_rng(seed), requests)
    say(f"serving budget: seed {seed}, {requests} requests")
    with lb.Server(0, budget, 4) as srv:
        for kind, m, src, n, s, p in plan:
            srv.submit(kind, m, src, n, seed=s, param=p)
        srv.drain()
        done = srv.poll()
        st = srv.stats()
    assert len(done) == requests
    assert st["evictions"] > 0 and st["peak_bytes"] <= budget, st


def test_serving_plan_every_request_matches_oracle(ctx, lb, ora):
    """The serving workload at a reduced request count: the benchmark's size
    law (log-uniform up to 4,194,304 triangles) and 70/30 refit/rebuild mix,
    all requests in the queue at once over 4 slots, and a 3 GiB budget that
    forces evictions.  Every completed build and refit is checked through its
    digests against a host replay (the uniform generator, the jitter, the
    oracle's build or refit with the mesh's topology); a refit may find its
    mesh evicted (and then every later refit of it must too).  Meshes still
    resident at the end are compared byte for byte."""
    requests = 240
    budget = 3 << 30
    plan = lb.wl_serving_plan(np.random.default_rng(5), requests)
    say(f"serving plan: {requests} requests, budget {budget >> 30} GiB")
    with lb.Server(0, budget, 4) as srv:
        srv.set_digest(True)
        ids = [srv.submit(kind, m, src, n, seed=seed, param=p)
               for kind, m, src, n, seed, p in plan]
        srv.drain()
        done = {c["request_id"]: c for c in srv.poll()}
        st = srv.stats()
        resident = {}
        for m in sorted({m for kind, m, _, _, _, _ in plan if kind == lb.REQ_BUILD}):
            try:
                resident[m] = srv.read_mesh(ctx, m)
            except lb.LBVHError:
                pass
    assert len(done) == requests
    assert st["evictions"] > 0 and st["peak_bytes"] <= budget
    say(f"serving plan: {st['evictions']} evictions, {len(resident)} meshes resident; replay")
    last = {}
    for i, (kind, m, _, _, _, _) in enumerate(plan):
        last[m] = i
    state, gone = {}, set()
    checked = {"build": 0, "refit": 0, "evicted": 0}
    for i, (rid, (kind, m, src, n, seed, p)) in enumerate(zip(ids, plan)):
        c = done[rid]
        if kind == lb.REQ_BUILD:
            assert c["status"] == 0, (i, c)
            tri = lb.wl_uniform(seed, n, p)
            wb, wo, wsk, wpm, _ = ora.build_mt(tri, None, ora.DIV_MUL_RCP)
            state[m] = (tri, wsk, wpm)
            checked["build"] += 1
        else:
            if m in gone or c["status"] == lb.ERR_NOT_FOUND:
                assert c["status"] == lb.ERR_NOT_FOUND, (i, c)
                gone.add(m)
                state.pop(m, None)
                checked["evicted"] += 1
                continue
            assert c["status"] == 0, (i, c)
            tri0, wsk, wpm = state[m]
            tri = lb.wl_jitter(seed, tri0, p)
            wb, wo = ora.refit_mt(tri, wsk, wpm)
            state[m] = (tri, wsk, wpm)
            checked["refit"] += 1
        assert (c["digest_bvh"], c["digest_tri"]) == ora.digest(wb, wo), (i, kind, m, n)
        if last[m] == i:
            if m in resident:
                assert np.array_equal(resident[m][0], wb), m
                assert np.array_equal(resident[m][1].view(np.uint32), wo.view(np.uint32)), m
            state.pop(m, None)
    say(f"serving plan: checked {checked}")
    assert checked["refit"] > 0 and checked["build"] > 0


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


def test_cancel_and_queueing_deadline(lb):
    """A request still in the queue can be cancelled: it completes with
    LBVH_ERR_CANCELLED and nothing of it runs (a cancelled build leaves no
    mesh).  With a maximum wait, requests not admitted in time complete with
    LBVH_ERR_TIMEOUT; without one, the same requests run."""
    with lb.Server(0, 0, 1) as srv:
        say("server: cancel queued requests")
        first = [srv.submit(lb.REQ_BUILD, m, lb.SRC_UNIFORM, 4000000, seed=m, param=0.01)
                 for m in range(4)]
        late = [srv.submit(lb.REQ_BUILD, 100 + m, lb.SRC_UNIFORM, 1000, seed=m, param=0.01)
                for m in range(8)]
        cancelled = []
        for r in late:   # one slot, two in flight: these almost surely still wait
            try:
                srv.cancel(r)
                cancelled.append(r)
            except lb.LBVHError as e:
                assert e.code in (lb.ERR_BUSY, lb.ERR_NOT_FOUND)
        done = wait_all(srv)
        assert cancelled
        assert all(done[r]["status"] == 0 for r in first)
        for m, r in enumerate(late):
            assert done[r]["status"] == (lb.ERR_CANCELLED if r in cancelled else 0)
            if r in cancelled:
                with pytest.raises(lb.LBVHError) as err:
                    srv.mesh(100 + m)
                assert err.value.code == lb.ERR_NOT_FOUND
        with pytest.raises(lb.LBVHError) as err:
            srv.cancel(first[0])  # completed
        assert err.value.code == lb.ERR_NOT_FOUND
        say("server: queueing deadline")
        srv.set_max_wait(1e-9)
        timed = [srv.submit(lb.REQ_BUILD, 200 + m, lb.SRC_UNIFORM, 1000, seed=m, param=0.01)
                 for m in range(6)]
        done = wait_all(srv)
        assert all(done[r]["status"] == lb.ERR_TIMEOUT for r in timed)
        srv.set_max_wait(0)
        again = [srv.submit(lb.REQ_BUILD, 200 + m, lb.SRC_UNIFORM, 1000, seed=m, param=0.01)
                 for m in range(6)]
        done = wait_all(srv)
        assert all(done[r]["status"] == 0 for r in again)
        st = srv.stats()
        assert st["submitted"] == st["completed"] == 4 + 8 + 6 + 6
