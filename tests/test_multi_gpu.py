"""GPU parity of the batched multi-source engines and the query server.

Every lane of a batch must equal the single-source oracle result for its
source, whatever else shares the batch; served queries must give the same
result alone and in a mix.
"""
import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu

INF = grus.INFINIT


def host_csr(g):
    from oracle import oracle_py as O
    off, col, w = g.csr()
    return O.CSR(off, col, w)


def sources(h, k, seed):
    deg = h.degrees()
    rng = np.random.default_rng(seed)
    nz = np.nonzero(deg)[0]
    s = [int(np.argmax(deg))] + [int(x) for x in rng.choice(nz, k - 3, replace=False)]
    s.append(int(np.nonzero(deg == 0)[0][0]) if (deg == 0).any() else 0)  # isolated
    s.append(s[1])  # duplicate source in two lanes
    return s


@pytest.mark.parametrize("k", [1, 7, 64])
def test_bfs_multi_rmat(ctx, oracle, k):
    with grus.Graph.rmat(ctx, 14) as g:
        h = host_csr(g)
        srcs = sources(h, max(k, 3), seed=k)[:k]
        out, st = grus.bfs_multi(g, srcs)
        for i, s in enumerate(srcs):
            assert (out[i] == oracle.bfs(h, s)).all(), (i, s)


def test_bfs_multi_lane_reuse_and_grid(ctx, oracle):
    R, C = 90, 70
    with grus.Graph.grid(ctx, R, C) as g:
        rr, cc = np.divmod(np.arange(R * C), C)
        srcs = [0, 5 * C + 7, R * C - 1]
        out, _ = grus.bfs_multi(g, srcs)
        for i, s in enumerate(srcs):
            sr, sc = divmod(s, C)
            assert (out[i] == np.abs(rr - sr) + np.abs(cc - sc)).all()
        # second batch on the same engine: lanes are reset before reuse
        out2, _ = grus.bfs_multi(g, [C + 1] * 2 + [17])
        assert (out2[2] == np.abs(rr - 0) + np.abs(cc - 17)).all()


def test_bfs_multi_directed(ctx, oracle):
    rng = np.random.default_rng(8)
    V = 4000
    u = rng.integers(0, V, 25000)
    v = (u + rng.integers(1, 40, 25000)) % V
    key = np.unique(u.astype(np.uint64) << np.uint64(32) | v.astype(np.uint64))
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        srcs = list(range(0, V, 97))[:40]
        out, _ = grus.bfs_multi(g, srcs)
        for i, s in enumerate(srcs):
            assert (out[i] == oracle.bfs(h, s)).all()


@pytest.mark.parametrize("k", [1, 64, 128])
def test_sssp_multi_rmat(ctx, oracle, k):
    with grus.Graph.rmat(ctx, 12, weights=(1, 255), wseed=3) as g:
        h = host_csr(g)
        srcs = sources(h, max(k, 3), seed=3 + k)[:k]
        out, st = grus.sssp_multi(g, srcs)
        want = 0
        for i, s in enumerate(srcs):
            ref = oracle.sssp(h, s)
            assert (out[i] == ref).all(), (i, s)
            want += int((ref != INF).sum())
        # the reached count sums every slot's, copied out or not
        assert st.vertices_reached == want
        assert grus.sssp_multi(g, srcs, copy=False)[1].vertices_reached == want


def test_bfs_multi_reached_count(ctx, oracle):
    with grus.Graph.rmat(ctx, 12) as g:
        h = host_csr(g)
        srcs = sources(h, 40, seed=11)
        want = sum(int((oracle.bfs(h, s) != INF).sum()) for s in srcs)
        assert grus.bfs_multi(g, srcs)[1].vertices_reached == want
        assert grus.bfs_multi(g, srcs, copy=False)[1].vertices_reached == want


def test_sssp_multi_128_sources_and_errors(ctx, oracle):
    with grus.Graph.rmat(ctx, 11, weights=(1, 255), wseed=9) as g:
        h = host_csr(g)
        rng = np.random.default_rng(3)
        srcs = [int(x) for x in rng.integers(0, g.V, 128)]
        out, _ = grus.sssp_multi(g, srcs)
        for i in range(0, 128, 9):
            assert (out[i] == oracle.sssp(h, srcs[i])).all(), i
        with pytest.raises(grus.GrusError) as e:
            grus.sssp_multi(g, list(range(129)))
        assert e.value.code == capi.E_INVALID
        with pytest.raises(grus.GrusError) as e:
            grus.bfs_multi(g, list(range(65)))
        assert e.value.code == capi.E_INVALID


def test_sssp_multi_grid(ctx, oracle):
    with grus.Graph.grid(ctx, 60, 80, weights=(1, 1000), wseed=2) as g:
        h = host_csr(g)
        srcs = [0, 1234, 4799, 2400]
        out, _ = grus.sssp_multi(g, srcs)
        for i, s in enumerate(srcs):
            assert (out[i] == oracle.sssp(h, s)).all()


def test_sssp_multi_distances_past_16_bits(ctx, oracle):
    # a heavy path (distances up to 299,000) beside a light random component:
    # rows start 16 bits wide, the path's lanes overflow, every row widens to
    # 32 bits and those lanes restart; every lane must still equal Dijkstra,
    # in this batch and in the next one on the widened engine
    rng = np.random.default_rng(21)
    P, Q = 300, 2000
    edges = [(i, i + 1, 1000) for i in range(P - 1)]
    for _ in range(8000):
        a, b = rng.integers(P, P + Q, 2)
        if a != b:
            edges.append((int(a), int(b), int(rng.integers(1, 11))))
    d = {}
    for a, b, w in edges:
        for x, y in ((a, b), (b, a)):
            d[(x, y)] = min(w, d.get((x, y), w))
    keys = sorted(d)
    V = P + Q
    off = np.zeros(V + 1, np.uint64)
    for a, _ in keys:
        off[a + 1] += 1
    off = np.cumsum(off).astype(np.uint64)
    col = np.array([b for _, b in keys], np.uint32)
    w = np.array([d[k] for k in keys], np.uint32)
    h = oracle.CSR(off, col, w)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True) as g:
        # slots pair up in lanes (0,1), (2,3), ...: heavy and light sources
        # share lanes, so a lane keeps running while one of its slots restarts
        for srcs in ([P + 5, 0, P + 77, 150, P + 1999, 299], [P + 3, 1, P + 600]):
            out, _ = grus.sssp_multi(g, srcs)
            for i, s in enumerate(srcs):
                assert (out[i] == oracle.sssp(h, s)).all(), (i, s)
        assert int(oracle.sssp(h, 0)[P - 1]) == 299000


# ------------------------------------------------------------------ serving
def single_digests(g, kind, srcs):
    d = {}
    for s in set(srcs):
        (grus.bfs if kind == grus.Q_BFS else grus.sssp)(g, s, copy=False)
        d[s] = g.digest()
    return d


@pytest.mark.parametrize("opts", [
    {},  # default: one engine per kind, batches admitted into idle engines
    {"max_inflight_batches": 3, "admit_min_free": 1},  # lane-by-lane admission
    {"max_inflight_batches": 2, "max_batch": 16, "admit_min_free": 5},
], ids=["idle-only", "any-free-lane", "gate5of16"])
def test_server_mix_matches_single_queries(ctx, oracle, opts):
    with grus.Graph.rmat(ctx, 13, weights=(1, 255), wseed=3) as g:
        h = host_csr(g)
        rng = np.random.default_rng(11)
        nz = np.nonzero(h.degrees())[0]
        qs = [(int(rng.integers(0, 2)), int(rng.choice(nz))) for _ in range(300)]
        ref = {grus.Q_BFS: single_digests(g, grus.Q_BFS, [s for k, s in qs if k == 0]),
               grus.Q_SSSP: single_digests(g, grus.Q_SSSP, [s for k, s in qs if k == 1])}
        outs = {}
        with grus.Server(g, copy_results=True, **opts) as srv:
            ids = {}
            for kind, s in qs:
                buf = np.zeros(g.V, np.uint32)
                qid = srv.submit(kind, s, buf)
                ids[qid] = (kind, s)
                outs[qid] = buf
            got = {}
            for _ in range(len(qs)):
                r = srv.poll(60000)
                assert r.status == 0
                got[r.id] = r
        assert set(got) == set(ids)
        for qid, (kind, s) in ids.items():
            r = got[qid]
            assert (r.reached, r.checksum) == ref[kind][s], (kind, s)
            assert r.submit_us <= r.start_us <= r.done_us
        # full results of a sample against the oracle
        for qid in list(ids)[:12]:
            kind, s = ids[qid]
            exp = oracle.bfs(h, s) if kind == grus.Q_BFS else oracle.sssp(h, s)
            assert (outs[qid] == exp).all()


def test_server_alone_equals_in_mix(ctx):
    with grus.Graph.rmat(ctx, 12, weights=(1, 255), wseed=5) as g:
        with grus.Server(g) as srv:
            srv.submit(grus.Q_SSSP, 7)
            alone = srv.poll(60000)
            for s in range(100, 160):
                srv.submit(grus.Q_SSSP, s)
            srv.submit(grus.Q_SSSP, 7)
            for s in range(200, 230):
                srv.submit(grus.Q_BFS, s)
            res = [srv.poll(60000) for _ in range(91)]
        mixed = [r for r in res if r.id == 62][0]
        assert (mixed.reached, mixed.checksum) == (alone.reached, alone.checksum)


def _mixed_queries(g, n, seed):
    h = host_csr(g)
    rng = np.random.default_rng(seed)
    nz = np.nonzero(h.degrees())[0]
    return [(int(rng.integers(0, 2)), int(rng.choice(nz))) for _ in range(n)]


def _serve_all(srv, qs, cap=None):
    """Submits every query at once, polls them all; returns the results by
    query and the largest engine count seen (checking the state cap)."""
    ids = {srv.submit(kind, s): (kind, s) for kind, s in qs}
    got, most = {}, 0
    for _ in range(len(qs)):
        r = srv.poll(60000)
        assert r.status == 0
        got[r.id] = r
        st = srv.stats()
        if cap is not None:
            assert st["state_bytes"] <= cap, st
        most = max(most, sum(st["engines_held"]))
    assert set(got) == set(ids)
    return {ids[q]: (r.reached, r.checksum) for q, r in got.items()}, most


def test_server_admits_by_state_memory(ctx):
    # engines beyond the first of each kind take device state only while it
    # fits the state budget; the rest of a burst waits for it, and an idle
    # extra engine gives its state back
    import time
    with grus.Graph.rmat(ctx, 13, weights=(1, 255), wseed=3) as g:
        qs = _mixed_queries(g, 300, 17)
        ref = {grus.Q_BFS: single_digests(g, grus.Q_BFS, [s for k, s in qs if k == 0]),
               grus.Q_SSSP: single_digests(g, grus.Q_SSSP, [s for k, s in qs if k == 1])}
        with grus.Server(g, solo_max=-1, max_inflight_batches=1) as probe:
            st = probe.stats()
        assert st["engines_held"] == (1, 1) and st["engines_started"] == 2
        both = st["state_bytes"]  # one engine of each kind, about the same size
        cap = both * 8 // 5  # room for one more engine, not two
        with grus.Server(g, state_budget_bytes=cap, idle_release_ms=100, solo_max=-1) as srv:
            res, most = _serve_all(srv, qs, cap)
            st = srv.stats()
            assert most == 3, most
            assert st["memory_waits"] > 0 and st["engines_started"] >= 3, st
            time.sleep(0.6)  # past the idle time: the extra engine gives its state back
            st = srv.stats()
            assert st["engines_held"] == (1, 1) and st["engines_released"] >= 1, st
            assert st["state_bytes"] == both
            # state is taken again for the next burst
            res2, _ = _serve_all(srv, qs[:120], cap)
        for (kind, s), d in list(res.items()) + list(res2.items()):
            assert d == ref[kind][s], (kind, s)


def test_server_admits_by_device_budget(ctx):
    # the same with the device budget as the limit: an extra engine's
    # allocation is refused with GRUS_E_BUDGET and its batch waits
    with grus.Graph.rmat(ctx, 13, weights=(1, 255), wseed=3) as g:
        qs = _mixed_queries(g, 200, 23)
        ref = {grus.Q_BFS: single_digests(g, grus.Q_BFS, [s for k, s in qs if k == 0]),
               grus.Q_SSSP: single_digests(g, grus.Q_SSSP, [s for k, s in qs if k == 1])}
        held0, _ = ctx.device_memory()
        with grus.Server(g, solo_max=-1, max_inflight_batches=1) as probe:
            held1, _ = ctx.device_memory()
            both = probe.stats()["state_bytes"]
        budget = held1 + both // 4  # no room for another engine
        try:
            ctx.set_device_budget(budget)
            ctx.device_memory(reset_peak=True)
            with grus.Server(g, solo_max=-1) as srv:
                res, most = _serve_all(srv, qs)
                st = srv.stats()
            assert most == 2 and st["memory_waits"] > 0, st
            assert ctx.device_memory()[1] <= budget
        finally:
            ctx.set_device_budget(0)
        assert held1 > held0
        for (kind, s), d in res.items():
            assert d == ref[kind][s], (kind, s)


def test_server_uncapped_engines_hold_state_from_creation(ctx):
    with grus.Graph.rmat(ctx, 11) as g:
        with grus.Server(g, max_inflight_batches=3, solo_max=-1) as srv:
            st = srv.stats()
            assert st["engines_held"] == (3, 3) and st["engines_started"] == 6
            res, most = _serve_all(srv, _mixed_queries(g, 40, 3))
            assert most == 6 and srv.stats()["engines_released"] == 0


def test_server_errors(ctx):
    with grus.Graph.rmat(ctx, 10) as g:
        with pytest.raises(grus.GrusError) as e:
            grus.Server(g, state_budget_bytes=1024)
        assert e.value.code == capi.E_BUDGET
        for bad in ({"admit_min_free": -1}, {"max_batch": 8, "admit_min_free": 9}):
            with pytest.raises(grus.GrusError) as e:
                grus.Server(g, **bad)
            assert e.value.code == capi.E_INVALID
        with grus.Server(g, max_batch=4) as srv:
            with pytest.raises(grus.GrusError) as e:
                srv.submit(grus.Q_BFS, g.V + 5)
            assert e.value.code == capi.E_INVALID
            with pytest.raises(grus.GrusError) as e:
                srv.poll(10)
            assert e.value.code == capi.E_TIMEOUT
            for s in range(20):  # more queries than lanes: admitted as lanes free up
                srv.submit(grus.Q_BFS, s)
            res = [srv.poll(60000) for _ in range(20)]
            assert all(r.status == 0 for r in res)
        # destroy with queries still queued must not hang or crash
        srv = grus.Server(g, max_batch=2)
        for s in range(50):
            srv.submit(grus.Q_SSSP if s % 2 else grus.Q_BFS, s)
        srv.close()
