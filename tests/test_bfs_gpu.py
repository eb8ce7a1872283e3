"""GPU parity: device generators and BFS levels, bit-exact against the oracle."""
import json
import os

import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
INF = grus.INFINIT


def host_csr(g):
    from oracle import oracle_py as O
    off, col, w = g.csr()
    return O.CSR(off, col, w)


def gen_gpu(ctx, spec):
    if spec["kind"] == "rmat":
        w = spec.get("weights")
        return grus.Graph.rmat(ctx, spec["scale"], weights=tuple(w) if w else None,
                               wseed=spec.get("wseed", 0))
    w = spec.get("weights")
    return grus.Graph.grid(ctx, spec["rows"], spec["cols"], weights=tuple(w) if w else None,
                           wseed=spec.get("wseed", 0))


def test_device_generators_match_pins(ctx):
    with open(os.path.join(HERE, "golden", "generators.json")) as f:
        gold = json.load(f)
    for key, spec in gold.items():
        print("generator", key, flush=True)
        with gen_gpu(ctx, spec) as g:
            h = host_csr(g)
            assert h.E == spec["E"], key
            assert "%016x" % h.fnv() == spec["fnv"], key


def test_device_rmat_matches_oracle_generator(ctx, oracle):
    ref = oracle.gen_rmat(13, weights=(1, 255), wseed=9, edge_factor=8, seed=77)
    with grus.Graph.rmat(ctx, 13, edge_factor=8, seed=77, weights=(1, 255), wseed=9) as g:
        h = host_csr(g)
        assert (h.off == ref.off).all() and (h.col == ref.col).all() and (h.w == ref.w).all()
        info = g.info
        deg = ref.degrees()
        assert info.max_degree == deg.max()
        assert info.max_degree_vertex == int(np.argmax(deg))
        assert info.isolated == int((deg == 0).sum())


def sources(h, n_random=3, seed=0):
    deg = h.degrees()
    rng = np.random.default_rng(seed)
    nz = np.nonzero(deg)[0]
    s = [int(np.argmax(deg))] + [int(x) for x in rng.choice(nz, n_random, replace=False)]
    z = np.nonzero(deg == 0)[0]
    if len(z):
        s.append(int(z[0]))  # isolated source: only itself reached
    return s


@pytest.mark.parametrize("scale", [10, 16])
@pytest.mark.parametrize("direction", ["auto", "push", "pull"])
def test_bfs_rmat_vs_oracle(ctx, oracle, scale, direction):
    with grus.Graph.rmat(ctx, scale) as g:
        h = host_csr(g)
        for s in sources(h):
            print("bfs rmat", scale, direction, "src", s, flush=True)
            lv, st = grus.bfs(g, s, direction)
            ref = oracle.bfs(h, s)
            assert (lv == ref).all(), (scale, direction, s)
            reached = ref != INF
            assert st.vertices_reached == reached.sum()
            assert st.edges_reached == h.degrees()[reached].sum()
            assert st.iterations == (ref[reached].max() + 1 if reached.any() else 0)


def test_bfs_grid_high_diameter(ctx, oracle):
    R, C = 300, 200
    with grus.Graph.grid(ctx, R, C) as g:
        rr, cc = np.divmod(np.arange(R * C), C)
        for direction in ("auto", "push", "pull"):
            lv, st = grus.bfs(g, 0, direction)
            assert (lv == rr + cc).all(), direction
            assert st.iterations == R + C - 1


def test_bfs_directed_graph_all_directions(ctx, oracle):
    rng = np.random.default_rng(3)
    V = 5000
    u = rng.integers(0, V, 40000)
    v = (u + rng.integers(1, 50, 40000)) % V  # mostly forward edges, directed
    key = np.unique(u.astype(np.uint64) << np.uint64(32) | v.astype(np.uint64))
    key = key[(key >> np.uint64(32)) != (key & np.uint64(0xFFFFFFFF))]
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        for s in (0, 17, 4321):
            ref = oracle.bfs(h, s)
            for direction in ("auto", "push", "pull"):
                lv, _ = grus.bfs(g, s, direction)
                assert (lv == ref).all(), (s, direction)


def layered_graph_with_unreachable_block(rng):
    """Directed: source 0 (in-degree 0) -> layers of 150000, 20000, 5000,
    1000 vertices, each vertex with 2 parents in the layer above; an
    unreachable strongly connected block U at the lowest ids (1..2000)
    feeds 3 extra in-edges into half of the layer vertices (so their first
    in-neighbour is often in U, never in a frontier), and 60 layer-3
    vertices get 100 in-edges from U (rows past the bottom-up cut-off)."""
    nu, sizes = 2000, [150000, 20000, 5000, 1000]
    bounds = np.cumsum([1 + nu] + sizes)  # layer i holds ids [bounds[i], bounds[i + 1])
    V = int(bounds[-1])
    layer = [np.arange(bounds[i], bounds[i + 1]) for i in range(len(sizes))]
    U = np.arange(1, 1 + nu)
    us, vs = [np.zeros(sizes[0], np.int64)], [layer[0]]
    for i in range(1, len(sizes)):
        for _ in range(2):
            us.append(rng.choice(layer[i - 1], sizes[i]))
            vs.append(layer[i])
    for L in layer:
        half = rng.choice(L, len(L) // 2, replace=False)
        for _ in range(3):
            us.append(rng.choice(U, len(half)))
            vs.append(half)
    longv = rng.choice(layer[2], 60, replace=False)
    us.append(rng.choice(U, 60 * 100))
    vs.append(np.repeat(longv, 100))
    us.append(U)
    vs.append(np.roll(U, -1))  # a cycle through U
    us.append(rng.choice(U, 8000))
    vs.append(rng.choice(U, 8000))
    u = np.concatenate(us).astype(np.uint64)
    v = np.concatenate(vs).astype(np.uint64)
    key = np.unique(u << np.uint64(32) | v)
    key = key[(key >> np.uint64(32)) != (key & np.uint64(0xFFFFFFFF))]
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    return np.cumsum(off).astype(np.uint64), col


def test_bfs_sparse_bottom_up_steps(ctx, oracle):
    # late bottom-up steps scan a list of the unvisited vertices; here that
    # list keeps the unreachable block, first tests miss into it and some
    # rows go to the wave-cooperative scan
    off, col = layered_graph_with_unreachable_block(np.random.default_rng(23))
    h = oracle.CSR(off, col)
    ref = oracle.bfs(h, 0)
    assert (ref[1:2001] == INF).all() and ref[ref != INF].max() == 4
    assert (ref != INF).sum() == 176001
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        for direction in ("pull", "auto"):
            lv, st = grus.bfs(g, 0, direction)
            assert (lv == ref).all(), direction
            assert st.vertices_reached == (ref != INF).sum()
        for s in (int(np.nonzero(ref == 1)[0][5]), 1):  # a layer-1 vertex; a vertex of U
            r = oracle.bfs(h, s)
            lv, _ = grus.bfs(g, s, "pull")
            assert (lv == r).all(), s


def test_bfs_closed_forms_and_edge_cases(ctx):
    # path: level = index
    n = 3000
    off = np.concatenate([[0], np.cumsum([1] + [2] * (n - 2) + [1])]).astype(np.uint64)
    col = []
    for i in range(n):
        if i > 0:
            col.append(i - 1)
        if i < n - 1:
            col.append(i + 1)
    col = np.array(col, np.uint32)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        for d in ("auto", "push", "pull"):
            lv, _ = grus.bfs(g, 0, d)
            assert (lv == np.arange(n)).all()
        with pytest.raises(grus.GrusError) as e:
            grus.bfs(g, n, "auto")
        assert e.value.code == capi.E_INVALID
    # graph with no edges
    with grus.Graph.from_csr(ctx, np.zeros(11, np.uint64), np.zeros(0, np.uint32),
                              symmetric=True) as g:
        lv, st = grus.bfs(g, 4)
        exp = np.full(10, INF, np.uint32)
        exp[4] = 0
        assert (lv == exp).all() and st.vertices_reached == 1
    # single vertex
    with grus.Graph.from_csr(ctx, np.zeros(2, np.uint64), np.zeros(0, np.uint32),
                              symmetric=True) as g:
        lv, _ = grus.bfs(g, 0)
        assert lv.tolist() == [0]


def test_bfs_small_levels_hand_back(ctx, oracle):
    # a complete binary tree (frontiers double: levels run in one workgroup
    # until the frontier outgrows it, then return to the host loop) with a
    # 300-vertex path hung from its last leaf (back to small levels)
    n, tail = (1 << 15) - 1, 300
    child = np.arange(1, n, dtype=np.int64)
    src = np.concatenate([(child - 1) // 2, child])
    dst = np.concatenate([child, (child - 1) // 2])
    p = np.arange(n, n + tail, dtype=np.int64)
    prev = np.concatenate([[n - 1], p[:-1]])
    src = np.concatenate([src, prev, p])
    dst = np.concatenate([dst, p, prev])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    V = n + tail
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    col = dst.astype(np.uint32)
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        for s0 in (0, n - 1, V - 1):
            want = oracle.bfs(h, s0)
            for d in ("auto", "push"):
                lv, st = grus.bfs(g, s0, d)
                assert (lv == want).all(), (s0, d)
                assert st.iterations == int(want.max()) + 1
                assert st.vertices_reached == V
                assert st.edges_reached == int(off[-1])


def test_bfs_repeatable_and_result_copy(ctx):
    with grus.Graph.rmat(ctx, 14) as g:
        s = g.info.max_degree_vertex
        a, _ = grus.bfs(g, s)
        b, _ = grus.bfs(g, s)
        assert (a == b).all()
        _, _ = grus.bfs(g, s, copy=False)
        assert (g.result() == a).all()


def test_bfs_large_rmat_vs_oracle(ctx, oracle):
    # s20 on the device generator's CSR; oracle runs on the copied CSR
    with grus.Graph.rmat(ctx, 20) as g:
        h = host_csr(g)
        for s in sources(h, n_random=2, seed=1)[:3]:
            print("bfs s20 src", s, flush=True)
            lv, st = grus.bfs(g, s)
            assert (lv == oracle.bfs(h, s)).all()
            assert st.pull_iterations > 0  # the direction switch engaged
