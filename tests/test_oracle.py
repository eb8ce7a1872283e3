"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

The reference ships no tests or golden vectors, so each oracle function is
pinned by closed forms, brute force / independent library routines (scipy's
csgraph), and invariants, each chosen so that a plausible mistake (off-by-
one level, dropped weight, wrong degree in PageRank, transposed edges)
fails at least one check.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
INF = 0xFFFFFFFE


def csr_from_edges(V, edges, weights=None, symmetric=True):
    """CSR with sorted, deduplicated rows from an edge list (test helper)."""
    d = {}
    for i, (u, v) in enumerate(edges):
        w = 1 if weights is None else weights[i]
        pairs = [(u, v), (v, u)] if symmetric else [(u, v)]
        for a, b in pairs:
            if a != b:
                d[(a, b)] = min(w, d.get((a, b), w))
    keys = sorted(d)
    off = np.zeros(V + 1, np.uint64)
    for a, _ in keys:
        off[a + 1] += 1
    off = np.cumsum(off).astype(np.uint64)
    col = np.array([b for _, b in keys], np.uint32)
    w = np.array([d[k] for k in keys], np.uint32)
    return off, col, w


def scipy_matrix(g, weighted=False):
    data = g.w.astype(np.float64) if (weighted and g.w is not None) else np.ones(g.E)
    return sp.csr_matrix((data, g.col.astype(np.int64), g.off.astype(np.int64)),
                         shape=(g.V, g.V))


def as_u32_dist(d):
    out = np.full(len(d), INF, np.uint32)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint64).astype(np.uint32)
    return out


# ------------------------------------------------------------------ BFS
def test_bfs_closed_forms(oracle):
    n = 50
    # path P_n: level = |i - src|
    g = oracle.CSR(*csr_from_edges(n, [(i, i + 1) for i in range(n - 1)])[:2])
    for s in (0, 7, n - 1):
        assert (oracle.bfs(g, s) == np.abs(np.arange(n) - s)).all()
    # cycle C_n: min(i, n - i)
    g = oracle.CSR(*csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])[:2])
    i = np.arange(n)
    assert (oracle.bfs(g, 0) == np.minimum(i, n - i)).all()
    # star: centre 0
    g = oracle.CSR(*csr_from_edges(n, [(0, i) for i in range(1, n)])[:2])
    lv = oracle.bfs(g, 0)
    assert lv[0] == 0 and (lv[1:] == 1).all()
    lv = oracle.bfs(g, 5)
    assert lv[5] == 0 and lv[0] == 1 and (np.delete(lv, [0, 5]) == 2).all()
    # complete graph
    g = oracle.CSR(*csr_from_edges(12, [(a, b) for a in range(12) for b in range(a)])[:2])
    lv = oracle.bfs(g, 3)
    assert lv[3] == 0 and (np.delete(lv, 3) == 1).all()
    # grid from the corner: r + c
    R, C = 9, 13
    g = oracle.gen_grid(R, C)
    rr, cc = np.divmod(np.arange(R * C), C)
    assert (oracle.bfs(g, 0) == rr + cc).all()
    # disconnected vertex stays INFINIT
    g = oracle.CSR(*csr_from_edges(4, [(0, 1), (1, 2)])[:2])
    assert oracle.bfs(g, 0).tolist() == [0, 1, 2, INF]


@pytest.mark.parametrize("scale", [8, 11])
def test_bfs_vs_scipy(oracle, scale):
    g = oracle.gen_rmat(scale)
    M = scipy_matrix(g)
    for s in (int(np.argmax(g.degrees())), 1, 17):
        d = csgraph.shortest_path(M, indices=s, unweighted=True, directed=True)
        assert (oracle.bfs(g, s) == as_u32_dist(d)).all()


def test_bfs_directed_vs_scipy(oracle):
    rng = np.random.default_rng(5)
    V = 300
    edges = [tuple(x) for x in rng.integers(0, V, size=(1200, 2))]
    off, col, _ = csr_from_edges(V, edges, symmetric=False)
    g = oracle.CSR(off, col)
    d = csgraph.shortest_path(scipy_matrix(g), indices=4, unweighted=True, directed=True)
    assert (oracle.bfs(g, 4) == as_u32_dist(d)).all()


# ------------------------------------------------------------------ SSSP
def test_sssp_unit_grid_is_manhattan(oracle):
    R, C = 17, 23
    g = oracle.gen_grid(R, C)
    rr, cc = np.divmod(np.arange(R * C), C)
    assert (oracle.sssp(g, 0, unit=True) == rr + cc).all()
    s = 5 * C + 11
    assert (oracle.sssp(g, s, unit=True) == np.abs(rr - 5) + np.abs(cc - 11)).all()


def floyd_warshall(V, off, col, w):
    D = np.full((V, V), np.inf)
    np.fill_diagonal(D, 0)
    for u in range(V):
        for e in range(int(off[u]), int(off[u + 1])):
            D[u, col[e]] = min(D[u, col[e]], w[e])
    for k in range(V):
        D = np.minimum(D, D[:, k:k + 1] + D[k:k + 1, :])
    return D


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sssp_floyd_warshall(oracle, seed):
    rng = np.random.default_rng(seed)
    V = 48
    edges = [tuple(x) for x in rng.integers(0, V, size=(150, 2))]
    wts = rng.integers(1, 100, size=150)
    off, col, w = csr_from_edges(V, edges, wts, symmetric=bool(seed % 2))
    g = oracle.CSR(off, col, w)
    D = floyd_warshall(V, off, col, w)
    for s in (0, 9, 31):
        assert (oracle.sssp(g, s) == as_u32_dist(D[s])).all()


def test_sssp_weighted_rmat_vs_scipy_and_certificate(oracle):
    g = oracle.gen_rmat(10, weights=(1, 255), wseed=3)
    s = int(np.argmax(g.degrees()))
    d = oracle.sssp(g, s)
    ref = csgraph.dijkstra(scipy_matrix(g, weighted=True), indices=s)
    assert (d == as_u32_dist(ref)).all()
    # certificate: d[v] <= d[u] + w for every edge, equality on some in-edge
    u = np.repeat(np.arange(g.V), g.degrees())
    fin = d[u] != INF
    assert (d[g.col[fin]].astype(np.int64) <= d[u[fin]].astype(np.int64) + g.w[fin]).all()


def test_sssp_grid_weights_vs_scipy(oracle):
    g = oracle.gen_grid(40, 50, weights=(1, 1000), wseed=2)
    ref = csgraph.dijkstra(scipy_matrix(g, weighted=True), indices=0)
    assert (oracle.sssp(g, 0) == as_u32_dist(ref)).all()


# ------------------------------------------------------------------ CC
def min_label_scipy(g):
    n, lab = csgraph.connected_components(scipy_matrix(g), directed=False)
    mins = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(mins, lab, np.arange(g.V))
    return mins[lab].astype(np.uint32)


@pytest.mark.parametrize("scale", [6, 12])
def test_cc_vs_scipy_and_invariants(oracle, scale):
    g = oracle.gen_rmat(scale)
    lab = oracle.cc(g)
    assert (lab == min_label_scipy(g)).all()
    u = np.repeat(np.arange(g.V), g.degrees())
    assert (lab[u] == lab[g.col]).all()
    assert (lab <= np.arange(g.V)).all()
    assert (lab[lab] == lab).all()


def test_cc_many_components(oracle):
    V = 100
    edges = [(i, i + 1) for i in range(0, 90, 2)] + [(50, 70), (70, 99)]
    g = oracle.CSR(*csr_from_edges(V, edges)[:2])
    assert (oracle.cc(g) == min_label_scipy(g)).all()


# ------------------------------------------------------------------ PageRank
def test_pagerank_regular_graph_closed_form(oracle):
    # k-regular, no isolated vertex: r_k = (1 - a^(k+1)) / V at every vertex
    n, a = 40, 0.85
    g = oracle.CSR(*csr_from_edges(n, [(i, (i + 1) % n) for i in range(n)])[:2])
    for it in (0, 1, 5, 20):
        r = oracle.pagerank(g, a, it)
        assert np.allclose(r, (1 - a ** (it + 1)) / n, rtol=1e-14, atol=0)


def test_pagerank_star_exact_iterates(oracle):
    # star with L leaves plus one isolated vertex; exact rational iterates
    L, a = 6, Fraction(17, 20)
    V = L + 2
    g = oracle.CSR(*csr_from_edges(V, [(0, i) for i in range(1, L + 1)])[:2])
    b = (1 - a) / V
    c = leaf = b
    for it in range(1, 8):
        c, leaf = b + a * L * leaf, b + a * c / L
        r = oracle.pagerank(g, float(a), it)
        assert abs(r[0] - float(c)) < 1e-15
        assert np.allclose(r[1:L + 1], float(leaf), rtol=1e-14, atol=0)
        assert abs(r[V - 1] - float(b)) < 1e-16  # isolated: teleport only


def test_pagerank_directed_vs_dense(oracle):
    # independent dense formulation r' = b + a * A^T D^-1 r on a directed graph
    rng = np.random.default_rng(7)
    V = 60
    edges = [tuple(x) for x in rng.integers(0, V, size=(300, 2))]
    off, col, _ = csr_from_edges(V, edges, symmetric=False)
    g = oracle.CSR(off, col)
    t = oracle.transpose(g)
    A = np.zeros((V, V))
    for u in range(V):
        for e in range(int(off[u]), int(off[u + 1])):
            A[u, col[e]] = 1.0
    deg = A.sum(1)
    P = np.divide(A, deg[:, None], out=np.zeros_like(A), where=deg[:, None] > 0)
    a, r = 0.85, np.full(V, 0.15 / V)
    for _ in range(20):
        r = 0.15 / V + a * (P.T @ r)
    assert np.allclose(oracle.pagerank(g, a, 20, csc=t), r, rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------------ transpose
def test_transpose_vs_scipy(oracle):
    rng = np.random.default_rng(11)
    V = 200
    edges = [tuple(x) for x in rng.integers(0, V, size=(900, 2))]
    wts = rng.integers(1, 50, size=900)
    off, col, w = csr_from_edges(V, edges, wts, symmetric=False)
    g = oracle.CSR(off, col, w)
    t = oracle.transpose(g)
    M = sp.csr_matrix((w.astype(np.int64), col.astype(np.int64), off.astype(np.int64)),
                      shape=(V, V)).T.tocsr()
    M.sort_indices()
    assert (t.off == M.indptr).all()
    assert (t.col == M.indices).all()
    assert (t.w == M.data).all()


# ------------------------------------------------------------------ .gr reader
GR_EXPECT = {
    "odd_weighted": (4, [[(1, 7), (2, 3)], [(2, 1)], [(3, 9)], [(0, 4)]], True),
    "even_weighted": (3, [[(1, 2), (2, 5)], [(2, 8)], [(0, 6)]], True),
    "odd_unweighted": (4, [[(1, 1)], [(2, 1)], [], [(0, 1)]], False),
    "isolated": (5, [[], [(3, 11)], [], [(1, 11)], []], True),
}


@pytest.mark.parametrize("name", sorted(GR_EXPECT))
def test_oracle_gr_reader_fixtures(oracle, name):
    V, rows, weighted = GR_EXPECT[name]
    g, ety = oracle.read_gr(os.path.join(GOLD, name + ".gr"))
    assert ety == (4 if weighted else 0)
    assert g.V == V
    assert g.off.tolist() == [0] + list(np.cumsum([len(r) for r in rows]))
    assert g.col.tolist() == [d for r in rows for d, _ in r]
    assert g.w.tolist() == [w for r in rows for _, w in r]


# ------------------------------------------------------------------ generators
def test_rmat_quadrant_statistics(oracle):
    # first-level quadrant frequencies approach a, b, c, d
    n = 400_000
    u, v = oracle.rmat_samples(20, 0, n)
    top_u, top_v = u >> 19, v >> 19
    f = [np.mean((top_u == i) & (top_v == j)) for i, j in ((0, 0), (0, 1), (1, 0), (1, 1))]
    assert np.allclose(f, [0.57, 0.19, 0.19, 0.05], atol=0.004)
    # and at a deep level (independent draws per level)
    b_u, b_v = (u >> 3) & 1, (v >> 3) & 1
    assert abs(np.mean((b_u == 0) & (b_v == 0)) - 0.57) < 0.004


@pytest.mark.parametrize("scale", [6, 10])
def test_rmat_structure(oracle, scale):
    g = oracle.gen_rmat(scale, weights=(1, 255), wseed=3)
    M = 16 << scale
    assert g.E <= 2 * M
    u = np.repeat(np.arange(g.V), g.degrees())
    assert (u != g.col).all()  # no self-loops
    key = u.astype(np.uint64) << np.uint64(32) | g.col.astype(np.uint64)
    assert (key[1:] > key[:-1]).all()  # rows sorted, no duplicates
    rev = g.col.astype(np.uint64) << np.uint64(32) | u.astype(np.uint64)
    order = np.argsort(rev)
    assert (np.sort(rev) == key).all()  # symmetric
    assert (g.w[order] == g.w).all()  # weights symmetric
    assert g.w.min() >= 1 and g.w.max() <= 255


def test_generator_fixture_hashes(oracle):
    with open(os.path.join(GOLD, "generators.json")) as f:
        gold = json.load(f)
    for key, spec in gold.items():
        if spec["kind"] == "rmat":
            g = oracle.gen_rmat(spec["scale"], weights=spec.get("weights"),
                                wseed=spec.get("wseed", 0))
        else:
            g = oracle.gen_grid(spec["rows"], spec["cols"], weights=spec.get("weights"),
                                wseed=spec.get("wseed", 0))
        assert g.E == spec["E"], key
        assert "%016x" % g.fnv() == spec["fnv"], key


# ---------------------------------------------------------------- baselines
# The OpenMP baselines (oracle/baseline_omp.c) are what bench.py's
# cpu_baseline leg times; they must compute the oracle's results, or the
# GPU/CPU ratios compare different work.
@pytest.fixture(scope="module")
def baseline_graphs(oracle):
    return [("rmat12", oracle.gen_rmat(12, weights=(1, 255), wseed=3)),
            ("rmat14", oracle.gen_rmat(14, weights=(1, 255), wseed=3)),
            ("grid300x200", oracle.gen_grid(300, 200, weights=(1, 1000), wseed=2))]


@pytest.mark.parametrize("threads", [1, 4])
def test_baselines_equal_oracle(oracle, baseline_graphs, threads):
    for name, g in baseline_graphs:
        deg = g.degrees()
        srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][len(np.nonzero(deg)[0]) // 3])]
        if name.startswith("grid"):
            srcs.append(0)
        for s in srcs:
            lv, _ = oracle.bl_bfs(g, s, threads=threads)
            assert (lv == oracle.bfs(g, s)).all(), (name, s)
            for delta in (64, 1000):
                d, _ = oracle.bl_sssp(g, s, delta=delta, threads=threads)
                assert (d == oracle.sssp(g, s)).all(), (name, s, delta)
        lb, _ = oracle.bl_cc(g, threads=threads)
        assert (lb == oracle.cc(g)).all(), name
        for iters in (2, 20):
            r, _ = oracle.bl_pagerank(g, 0.85, iters, threads=threads)
            assert r.dtype == np.float32
            assert np.abs(r.astype(np.float64) - oracle.pagerank(g, 0.85, iters)).sum() <= 1e-6, \
                (name, iters)


def test_baselines_on_edge_cases(oracle):
    # an isolated source, a graph without edges, and two components
    off, col, w = csr_from_edges(6, [(0, 1), (1, 2), (4, 5)], weights=[3, 4, 7])
    g = oracle.CSR(off, col, w)
    for s in (0, 3, 5):
        assert (oracle.bl_bfs(g, s)[0] == oracle.bfs(g, s)).all()
        assert (oracle.bl_sssp(g, s)[0] == oracle.sssp(g, s)).all()
    assert (oracle.bl_cc(g)[0] == oracle.cc(g)).all()
    assert list(oracle.bl_cc(g)[0]) == [0, 0, 0, 3, 4, 4]
    e = oracle.CSR(np.zeros(5, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    assert list(oracle.bl_bfs(e, 2)[0]) == [INF, INF, 0, INF]
    assert list(oracle.bl_cc(e)[0]) == [0, 1, 2, 3]


def test_bfs_certificate_accepts_distances_and_rejects_planted_mistakes(oracle):
    g = oracle.gen_rmat(12)
    A = sp.csr_matrix((np.ones(g.E), g.col.astype(np.int64), g.off.astype(np.int64)),
                      shape=(g.V, g.V))
    src = int(np.argmax(g.degrees()))
    lv = oracle.bfs(g, src)
    hop = csgraph.shortest_path(A, unweighted=True, indices=src)
    assert (np.where(np.isinf(hop), INF, hop).astype(np.uint32) == lv).all()
    assert oracle.bfs_certify(g, src, lv) == (0, 0)
    grid = oracle.gen_grid(40, 30)
    assert oracle.bfs_certify(grid, 0, oracle.bfs(grid, 0)) == (0, 0)
    reached = np.nonzero((lv != INF) & (lv > 1))[0]
    v = int(reached[len(reached) // 2])
    for delta in (+1, -1):  # one level too high or too low at one vertex
        bad = lv.copy()
        bad[v] = int(lv[v]) + delta
        assert oracle.bfs_certify(g, src, bad)[0] in (2, 3), delta
    bad = lv.copy()
    bad[v] = 0  # a second level-0 vertex
    assert oracle.bfs_certify(g, src, bad) == (1, v)
    other = int(np.nonzero(lv == 1)[0][0])
    assert oracle.bfs_certify(g, other, lv)[0] == 1  # the wrong source
    unreached = np.nonzero(lv == INF)[0]
    assert len(unreached) > 0  # RMAT leaves isolated vertices
    # an unreached vertex that is reached by mistake with some level: its
    # (absent) neighbours cannot be its parent
    bad = lv.copy()
    bad[int(unreached[0])] = 3
    assert oracle.bfs_certify(g, src, bad)[0] == 3
    # a reached vertex wrongly left unreached: its neighbours see it
    bad = lv.copy()
    bad[v] = INF
    assert oracle.bfs_certify(g, src, bad)[0] == 2
