"""GPU parity: SSSP distances and CC labels bit-exact, PageRank L1 <= 1e-6 vs FP64."""
import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu

INF = grus.INFINIT
PR_TOL = 1e-6  # L1 over normalised ranks vs the FP64 oracle


def host_csr(g):
    from oracle import oracle_py as O
    off, col, w = g.csr()
    return O.CSR(off, col, w)


def random_directed(rng, V, M, wmax=None):
    u = rng.integers(0, V, M)
    v = rng.integers(0, V, M)
    key = np.unique(u.astype(np.uint64) << np.uint64(32) | v.astype(np.uint64))
    key = key[(key >> np.uint64(32)) != (key & np.uint64(0xFFFFFFFF))]
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    w = None if wmax is None else rng.integers(1, wmax + 1, len(col)).astype(np.uint32)
    return off, col, w


# ------------------------------------------------------------------ SSSP
@pytest.mark.parametrize("scale", [10, 15])
def test_sssp_rmat_vs_oracle(ctx, oracle, scale):
    with grus.Graph.rmat(ctx, scale, weights=(1, 255), wseed=3) as g:
        h = host_csr(g)
        deg = h.degrees()
        srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][5]), int(np.nonzero(deg == 0)[0][0])]
        for s in srcs:
            print("sssp rmat", scale, "src", s, flush=True)
            d, st = grus.sssp(g, s)
            ref = oracle.sssp(h, s)
            assert (d == ref).all(), s
            reached = ref != INF
            assert st.vertices_reached == reached.sum()
            assert st.edges_reached == deg[reached].sum()


def test_sssp_grid_vs_oracle_and_delta_invariance(ctx, oracle):
    with grus.Graph.grid(ctx, 256, 320, weights=(1, 1000), wseed=2) as g:
        h = host_csr(g)
        ref = oracle.sssp(h, 0)
        d, _ = grus.sssp(g, 0)
        assert (d == ref).all()
        t0 = ctx.tuning
        try:
            for delta in (1, 333, 100000, 0xFFFFFFFF):
                t = ctx.tuning
                t.sssp_delta = delta
                ctx.tuning = t
                print("sssp grid delta", delta, flush=True)
                d, _ = grus.sssp(g, 0)
                assert (d == ref).all(), delta
        finally:
            ctx.tuning = t0
        s = 256 * 160 + 77
        assert (grus.sssp(g, s)[0] == oracle.sssp(h, s)).all()


def test_sssp_unit_weights_equal_bfs(ctx, oracle):
    with grus.Graph.rmat(ctx, 12) as g:  # unweighted: weight 1 per edge
        s = g.info.max_degree_vertex
        assert (grus.sssp(g, s)[0] == grus.bfs(g, s)[0]).all()
    R, C = 40, 70
    with grus.Graph.grid(ctx, R, C) as g:
        rr, cc = np.divmod(np.arange(R * C), C)
        assert (grus.sssp(g, 0)[0] == rr + cc).all()  # Manhattan distance


def test_sssp_directed_weighted(ctx, oracle):
    rng = np.random.default_rng(9)
    off, col, w = random_directed(rng, 3000, 20000, wmax=500)
    h = oracle.CSR(off, col, w)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=False) as g:
        for s in (0, 11, 2999):
            assert (grus.sssp(g, s)[0] == oracle.sssp(h, s)).all()


def test_sssp_repeatable(ctx):
    with grus.Graph.rmat(ctx, 14, weights=(1, 255), wseed=3) as g:
        s = g.info.max_degree_vertex
        a, _ = grus.sssp(g, s)
        b, _ = grus.sssp(g, s)
        assert (a == b).all()


# ------------------------------------------------------------------ PageRank
@pytest.mark.parametrize("scale", [10, 16])
def test_pagerank_rmat_vs_fp64(ctx, oracle, scale):
    with grus.Graph.rmat(ctx, scale) as g:
        h = host_csr(g)
        r, st = grus.pagerank(g, 0.85, 20)
        ref = oracle.pagerank(h, 0.85, 20)
        l1 = np.abs(r.astype(np.float64) - ref).sum()
        print("pagerank s%d L1 %.3e" % (scale, l1), flush=True)
        assert l1 <= PR_TOL
        assert st.iterations == 20


def test_pagerank_closed_forms(ctx):
    # cycle (2-regular): r_k = (1 - a^(k+1)) / V at every vertex
    n, a = 5000, 0.85
    off = np.arange(0, 2 * n + 1, 2, dtype=np.uint64)
    col = np.stack([(np.arange(n) - 1) % n, (np.arange(n) + 1) % n], 1)
    col = np.sort(col, 1).reshape(-1).astype(np.uint32)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        for it in (0, 1, 7, 20):
            r, _ = grus.pagerank(g, a, it)
            assert np.allclose(r, (1 - a ** (it + 1)) / n, rtol=2e-6, atol=0), it


def test_pagerank_hub_rows_span_windows(ctx, oracle):
    # a star with a 50000-leaf centre: the centre's row spans ~25 windows
    L = 50000
    V = L + 3
    rows = [list(range(1, L + 1))] + [[0]] * L + [[V - 1], [V - 2]]
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    col = np.concatenate([np.array(r, np.uint32) for r in rows])
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        r, _ = grus.pagerank(g, 0.85, 20)
        ref = oracle.pagerank(h, 0.85, 20)
        assert np.abs(r - ref).sum() <= PR_TOL
        assert abs(r[0] - ref[0]) / ref[0] < 1e-5


def test_pagerank_directed_uses_out_degree(ctx, oracle):
    rng = np.random.default_rng(4)
    off, col, _ = random_directed(rng, 4000, 30000)
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        r, _ = grus.pagerank(g, 0.85, 20)
        ref = oracle.pagerank(h, 0.85, 20, csc=oracle.transpose(h))
        assert np.abs(r - ref).sum() <= PR_TOL


def test_pagerank_bit_repeatable(ctx):
    with grus.Graph.rmat(ctx, 16) as g:
        a, _ = grus.pagerank(g)
        b, _ = grus.pagerank(g)
        assert (a.view(np.uint32) == b.view(np.uint32)).all()


# ------------------------------------------------------------------ CC
@pytest.mark.parametrize("scale", [10, 16])
@pytest.mark.parametrize("method", ["hook", "lp"])
def test_cc_rmat_vs_oracle(ctx, oracle, scale, method):
    with grus.Graph.rmat(ctx, scale) as g:
        h = host_csr(g)
        lab, _ = grus.cc(g, method)
        assert (lab == oracle.cc(h)).all()


@pytest.mark.parametrize("method", ["hook", "lp"])
def test_cc_many_components_and_grid(ctx, oracle, method):
    rng = np.random.default_rng(2)
    V = 20000
    # sparse random symmetric graph: many small components
    u = rng.integers(0, V, 8000)
    v = rng.integers(0, V, 8000)
    key = np.unique(np.concatenate([u.astype(np.uint64) << np.uint64(32) | v.astype(np.uint64),
                                    v.astype(np.uint64) << np.uint64(32) | u.astype(np.uint64)]))
    key = key[(key >> np.uint64(32)) != (key & np.uint64(0xFFFFFFFF))]
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        assert (grus.cc(g, method)[0] == oracle.cc(h)).all()
    with grus.Graph.grid(ctx, 100, 130) as g:
        assert (grus.cc(g, method)[0] == 0).all()


def test_cc_directed_rejected(ctx):
    off = np.array([0, 1, 1], np.uint64)
    col = np.array([1], np.uint32)
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        with pytest.raises(grus.GrusError) as e:
            grus.cc(g)
        assert e.value.code == capi.E_UNSUPPORTED


# ------------------------------------------------------------------ larger
def test_apps_rmat_s20_vs_oracle(ctx, oracle):
    with grus.Graph.rmat(ctx, 20, weights=(1, 255), wseed=3) as g:
        h = host_csr(g)
        s = g.info.max_degree_vertex
        print("s20 sssp", flush=True)
        assert (grus.sssp(g, s)[0] == oracle.sssp(h, s)).all()
        print("s20 cc", flush=True)
        assert (grus.cc(g)[0] == oracle.cc(h)).all()
        print("s20 pagerank", flush=True)
        r, _ = grus.pagerank(g)
        assert np.abs(r - oracle.pagerank(h)).sum() <= PR_TOL


def test_pagerank_relabelled_path(ctx, oracle):
    with grus.Graph.rmat(ctx, 14) as g:
        h = host_csr(g)
        t0 = ctx.tuning
        try:
            t = ctx.tuning
            t.pr_layout = 1
            ctx.tuning = t
            r, _ = grus.pagerank(g)
        finally:
            ctx.tuning = t0
        assert np.abs(r - oracle.pagerank(h)).sum() <= PR_TOL


def pagerank_with(ctx, g, layout, bits=0, iters=20, task_edges=0):
    t0 = ctx.tuning
    try:
        t = ctx.tuning
        t.pr_layout = layout
        t.pr_bin_bits = bits
        t.pr_task_edges = task_edges
        ctx.tuning = t
        return grus.pagerank(g, 0.85, iters)
    finally:
        ctx.tuning = t0


@pytest.mark.parametrize("bits", [6, 10, 14])
def test_pagerank_blocked_equals_pull(ctx, oracle, bits):
    # the blocked layout sums the same terms in the same order as the pull:
    # bit-identical ranks, for every bin width, symmetric and directed
    rng = np.random.default_rng(11)
    off, col, _ = random_directed(rng, 30000, 400000)
    with grus.Graph.rmat(ctx, 15) as g, \
            grus.Graph.from_csr(ctx, off, col, symmetric=False) as d:
        for gr in (g, d):
            a, _ = pagerank_with(ctx, gr, 2)
            b, st = pagerank_with(ctx, gr, 3, bits)
            print("bits %d V=%d launches pull/blocked %d" % (bits, gr.V, st.kernel_launches),
                  flush=True)
            assert (a.view(np.uint32) == b.view(np.uint32)).all()
            assert st.kernel_launches == 1 + 3 * 20  # init + (scatter, pull, fixup) per iteration
        r, _ = pagerank_with(ctx, g, 3, bits)
        assert np.abs(r.astype(np.float64) - oracle.pagerank(host_csr(g), 0.85, 20)).sum() <= PR_TOL


def test_pagerank_blocked_edge_cases(ctx, oracle):
    # isolated vertices, a hub spanning every bin, rows of one edge, V not a
    # multiple of the source chunk; iteration counts 0 and 1
    L = 70001
    V = L + 5
    rows = [list(range(1, L + 1))] + [[0]] * L + [[], [V - 2], [V - 3], []]
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    col = np.concatenate([np.array(r, np.uint32) for r in rows])
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        for it in (0, 1, 20):
            a, _ = pagerank_with(ctx, g, 2, iters=it)
            b, _ = pagerank_with(ctx, g, 3, 8, iters=it)
            assert (a.view(np.uint32) == b.view(np.uint32)).all(), it
        assert np.abs(b - oracle.pagerank(h, 0.85, 20)).sum() <= PR_TOL


def test_pagerank_ragged_in_rows(ctx, oracle):
    # in-edge rows built to hit every window case: runs of empty rows inside
    # a window, rows of 1/7/8/9 edges (thread boundaries), rows of exactly a
    # window (2048) and one edge off, rows ending on window boundaries
    rng = np.random.default_rng(21)
    lens = []
    for _ in range(40):
        lens += [0] * int(rng.integers(0, 3000))
        lens += list(rng.choice([1, 7, 8, 9, 2047, 2048, 2049, 4096, 31, 0],
                                int(rng.integers(1, 40))))
    lens = np.array(lens, np.int64)
    V = len(lens)
    in_off = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    in_col = np.concatenate([np.sort(rng.choice(V, l, replace=False)) if l else
                             np.zeros(0, np.int64) for l in lens]).astype(np.uint32)
    out = oracle.transpose(oracle.CSR(in_off, in_col))
    with grus.Graph.from_csr(ctx, out.off, out.col, symmetric=False) as g:
        ref = oracle.pagerank(out, 0.85, 20, csc=oracle.transpose(out))
        a, _ = pagerank_with(ctx, g, 2)
        b, _ = pagerank_with(ctx, g, 3, 10)
        print("ragged V=%d E=%d L1 %.3e" % (V, g.E, np.abs(a - ref).sum()), flush=True)
        assert np.abs(a.astype(np.float64) - ref).sum() <= PR_TOL
        assert (a.view(np.uint32) == b.view(np.uint32)).all()


# ------------------------------------------------ PageRank, binned push layout
@pytest.mark.parametrize("bits,task_edges", [(14, 0), (10, 256), (6, 5000), (14, 300)])
def test_pagerank_binned_vs_fp64(ctx, oracle, bits, task_edges):
    # symmetric RMAT and a directed graph; tiny task sizes split every large
    # bin over several tasks (sums meet in the global accumulator)
    rng = np.random.default_rng(12)
    off, col, _ = random_directed(rng, 30000, 400000)
    with grus.Graph.rmat(ctx, 15) as g, \
            grus.Graph.from_csr(ctx, off, col, symmetric=False) as d:
        for gr in (g, d):
            h = host_csr(gr)
            csc = None if gr is g else oracle.transpose(h)
            ref = oracle.pagerank(h, 0.85, 20, csc=csc)
            a, _ = pagerank_with(ctx, gr, 2)
            b, st = pagerank_with(ctx, gr, 4, bits, task_edges=task_edges)
            c, _ = pagerank_with(ctx, gr, 4, bits, task_edges=task_edges)
            l1 = np.abs(b.astype(np.float64) - ref).sum()
            print("binned bits %d task %d V=%d L1 %.3e (pull %.3e) launches %d"
                  % (bits, task_edges, gr.V, l1, np.abs(a.astype(np.float64) - ref).sum(),
                     st.kernel_launches), flush=True)
            assert l1 <= PR_TOL
            assert np.abs(a.astype(np.float64) - b).sum() <= PR_TOL
            assert (b.view(np.uint32) == c.view(np.uint32)).all()  # run to run
            if task_edges and task_edges < 1000:  # split bins: a finish pass per iteration
                assert st.kernel_launches == 1 + 3 * 20


def test_pagerank_binned_bits_and_tasks_do_not_change_ranks(ctx):
    # every edge adds the same fixed-point term to the same integer sum
    # whatever the bin width and task split: bit-identical ranks
    with grus.Graph.rmat(ctx, 15) as g:
        ref, _ = pagerank_with(ctx, g, 4, 14)
        for bits, te in ((13, 0), (10, 256), (6, 5000), (14, 300), (12, 1 << 18)):
            r, _ = pagerank_with(ctx, g, 4, bits, task_edges=te)
            assert (r.view(np.uint32) == ref.view(np.uint32)).all(), (bits, te)


@pytest.mark.skip(reason="HIP-graph capture of the PageRank loop withdrawn: in its one GPU run the "
                         "first BFS after a graph replay hung (no fault); cause not found, so "
                         "the loop launches eagerly and this case stays off the GPU")
def test_pagerank_binned_graph_replays(ctx):
    # the iteration loop is captured into a HIP graph and replayed; a
    # CU-masked context (mask = every CU) runs the same loop eagerly.  Every
    # replay, with other work and re-captures in between, must equal it.
    total = ctx.cu_count[1]
    every_cu = [0xFFFFFFFF] * ((total + 31) // 32)
    with grus.Graph.rmat(ctx, 14) as g, grus.Graph.rmat(ctx, 12) as h:
        ctx.set_cu_mask(every_cu)
        try:
            eager20, _ = pagerank_with(ctx, g, 4)
            eager7, _ = pagerank_with(ctx, g, 4, iters=7)
            eager_h, _ = pagerank_with(ctx, h, 4)
        finally:
            ctx.set_cu_mask(None)
        for order in ((20, 20, 20), (7, 7), (20,)):
            for it in order:
                r, st = pagerank_with(ctx, g, 4, iters=it)
                ref = eager20 if it == 20 else eager7
                assert (r.view(np.uint32) == ref.view(np.uint32)).all(), it
                grus.bfs(g, g.info.max_degree_vertex)  # eager work between replays
            r, _ = pagerank_with(ctx, h, 4)  # another graph's own capture
            assert (r.view(np.uint32) == eager_h.view(np.uint32)).all()


def test_pagerank_binned_edge_cases(ctx, oracle):
    # a hub whose out-edges reach every bin, isolated vertices, V not a
    # multiple of the bin width, iteration counts 0 / 1 / 20, split bins
    L = 70001
    V = L + 5
    rows = [list(range(1, L + 1))] + [[0]] * L + [[], [V - 2], [V - 3], []]
    off = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.uint64)
    col = np.concatenate([np.array(r, np.uint32) for r in rows])
    h = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=True) as g:
        for bits, te in ((8, 0), (8, 256), (14, 1000)):
            for it in (0, 1, 20):
                a, _ = pagerank_with(ctx, g, 2, iters=it)
                b, _ = pagerank_with(ctx, g, 4, bits, iters=it, task_edges=te)
                ref = oracle.pagerank(h, 0.85, it)
                assert np.abs(b.astype(np.float64) - ref).sum() <= PR_TOL, (bits, te, it)
                assert np.abs(a.astype(np.float64) - b).sum() <= PR_TOL, (bits, te, it)
                if it == 0:
                    assert (a.view(np.uint32) == b.view(np.uint32)).all()
    # one vertex, no edges: the pull handles it
    with grus.Graph.from_csr(ctx, np.zeros(2, np.uint64), np.zeros(0, np.uint32),
                             symmetric=True) as g:
        r, _ = pagerank_with(ctx, g, 4)
        assert np.allclose(r, [0.15], rtol=1e-6)


def test_pagerank_binned_falls_back_to_pull(ctx):
    # 2^20 vertices in bins of 2^6 is 16384 bins, more than the layout
    # takes: the plain pull runs and gives its bit-identical ranks
    with grus.Graph.rmat(ctx, 20) as g:
        a, sa = pagerank_with(ctx, g, 2, iters=3)
        b, sb = pagerank_with(ctx, g, 4, 6, iters=3)
        assert (a.view(np.uint32) == b.view(np.uint32)).all()
        assert sa.kernel_launches == sb.kernel_launches


def sssp_with(ctx, g, src, pull_div, layout=0):
    t0 = ctx.tuning
    try:
        t = ctx.tuning
        t.sssp_pull_div = pull_div
        t.sssp_pull_layout = layout
        ctx.tuning = t
        return grus.sssp(g, src)
    finally:
        ctx.tuning = t0


ALWAYS_PULL, NEVER_PULL = 0xFFFFFFFE, 0xFFFFFFFF


def test_sssp_pull_rounds_match_push(ctx, oracle):
    # every round pulled, every round pushed, and the automatic mix all give
    # Dijkstra's distances: symmetric weighted RMAT and a directed graph
    # (whose pull rounds read the transpose with its weights)
    rng = np.random.default_rng(8)
    off, col, w = random_directed(rng, 5000, 300000, wmax=1000)  # max degree > 32: host rounds
    with grus.Graph.rmat(ctx, 14, weights=(1, 255), wseed=3) as g, \
            grus.Graph.from_csr(ctx, off, col, w, symmetric=False) as d:
        for gr, src in ((g, g.info.max_degree_vertex), (d, 7)):
            ref = oracle.sssp(host_csr(gr), src)
            # pull layouts: 1 window over the in-edges, 2 binned push
            for div, layout in ((ALWAYS_PULL, 1), (ALWAYS_PULL, 2), (NEVER_PULL, 0), (32, 1),
                                (32, 2)):
                dist, st = sssp_with(ctx, gr, src, div, layout)
                print("V=%d div=%x layout %d rounds %d pull %d" % (
                    gr.V, div, layout, st.iterations, st.pull_iterations), flush=True)
                assert (dist == ref).all()
                if div == ALWAYS_PULL:  # a queue of out-degree-0 vertices still pushes
                    assert st.pull_iterations >= 1
                if div == NEVER_PULL:
                    assert st.pull_iterations == 0


def test_sssp_binned_rounds_layout_variants(ctx, oracle):
    # binned pull rounds over layouts of several shapes: bins of 2^8 / 2^12 /
    # 2^14 destinations, bins split over many tasks (their minima meet in a
    # device array that must be all-infinite again after every round), a
    # layout PageRank built first and SSSP reuses, a directed graph, and
    # isolated vertices; every variant equals Dijkstra
    # weight streams of 1, 2 and 4 bytes: largest weights 255, 1000, 300000
    rng = np.random.default_rng(21)
    off, col, w = random_directed(rng, 6000, 200000, wmax=255)
    off4, col4, w4 = random_directed(rng, 5000, 150000, wmax=300000)
    with grus.Graph.rmat(ctx, 14, weights=(1, 255), wseed=4) as g, \
            grus.Graph.rmat(ctx, 13, weights=(1, 1000), wseed=5) as g2, \
            grus.Graph.from_csr(ctx, off, col, w, symmetric=False) as d, \
            grus.Graph.from_csr(ctx, off4, col4, w4, symmetric=False) as d4:
        t0 = ctx.tuning
        try:
            for gr, srcs in ((g, (g.info.max_degree_vertex, 3)), (g2, (g2.info.max_degree_vertex,)),
                             (d, (7, 4000)), (d4, (1, 2500))):
                h = host_csr(gr)
                refs = {s: oracle.sssp(h, s) for s in srcs}
                for bits, task_edges in ((14, 0), (8, 256), (12, 1024)):
                    t = ctx.tuning
                    t.pr_layout = 4
                    t.pr_bin_bits = bits
                    t.pr_task_edges = task_edges
                    t.sssp_pull_div = ALWAYS_PULL
                    t.sssp_pull_layout = 2
                    ctx.tuning = t
                    grus.pagerank(gr, iters=1, copy=False)  # (re)builds the shared layout
                    for s in srcs:
                        dist, st = grus.sssp(gr, s)
                        print("V=%d bits %d tasks %d src %d rounds %d pull %d" % (
                            gr.V, bits, task_edges, s, st.iterations, st.pull_iterations),
                            flush=True)
                        assert st.pull_iterations >= 1
                        assert (dist == refs[s]).all()
        finally:
            ctx.tuning = t0


def test_tuning_rejects_out_of_range_layouts(ctx):
    t0 = ctx.tuning
    try:
        for field, bad in (("sssp_pull_layout", 3), ("pr_layout", 5), ("pr_bin_bits", 5),
                           ("pr_task_edges", 100)):
            t = ctx.tuning
            setattr(t, field, bad)
            with pytest.raises(grus.GrusError) as e:
                ctx.tuning = t
            assert e.value.code == capi.E_INVALID, field
        assert ctx.tuning.sssp_pull_layout == t0.sssp_pull_layout  # left unchanged
    finally:
        ctx.tuning = t0


def test_sssp_grid_local_hops_invariant(ctx, oracle):
    # the persistent kernel's in-workgroup hop passes only add relaxations:
    # distances equal Dijkstra's for any pass count
    with grus.Graph.grid(ctx, 96, 160, weights=(1, 1000), wseed=2) as g:
        ref = oracle.sssp(host_csr(g), 5)
        t0 = ctx.tuning
        try:
            for hops in (1, 8, 64, 256):
                t = ctx.tuning
                t.sssp_local_hops = hops
                ctx.tuning = t
                d, st = grus.sssp(g, 5)
                print("hops %d rounds %d" % (hops, st.iterations), flush=True)
                assert (d == ref).all()
        finally:
            ctx.tuning = t0


def _binary_tree(n, wmax, seed):
    """Symmetric CSR of the complete binary tree on n vertices (i's children
    2i+1, 2i+2), one random weight in [1, wmax] per tree edge."""
    rng = np.random.default_rng(seed)
    child = np.arange(1, n, dtype=np.int64)
    parent = (child - 1) // 2
    wt = rng.integers(1, wmax + 1, n - 1).astype(np.uint32)
    src = np.concatenate([parent, child])
    dst = np.concatenate([child, parent])
    ww = np.concatenate([wt, wt])
    order = np.lexsort((dst, src))
    src, dst, ww = src[order], dst[order], ww[order]
    off = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    return off, dst.astype(np.uint32), ww


def test_sssp_persistent_far_list_overflow(ctx, oracle):
    # unit weights and bucket width 1 on a binary tree of 2^20 - 1 vertices:
    # every improvement lands in the far list, and the round that relaxes
    # depth 18 (2^18 vertices spread over the workgroups) makes each
    # workgroup stage several thousand far vertices, past its LDS list, so
    # the overflow path (stamp and publish at once) runs beside the round-
    # end stamping of the staged ones; a far vertex is listed once per epoch
    off, col, w = _binary_tree((1 << 20) - 1, 1, 3)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True) as g:
        ref = oracle.sssp(host_csr(g), 0)
        t0 = ctx.tuning
        try:
            for delta in (1, 2):
                t = ctx.tuning
                t.sssp_delta = delta
                ctx.tuning = t
                d, st = grus.sssp(g, 0)
                print("delta %d rounds %d" % (delta, st.iterations), flush=True)
                assert (d == ref).all()
            # and from a leaf, where the far list fills from the bottom up
            leaf = (1 << 20) - 2
            d, _ = grus.sssp(g, leaf)
            assert (d == oracle.sssp(host_csr(g), leaf)).all()
        finally:
            ctx.tuning = t0


def test_sssp_persistent_empty_buckets(ctx, oracle):
    # weights up to 100,000 on a binary tree (the persistent path) with
    # buckets much narrower than most weights: most far splits find the next
    # bucket [T, T + delta) empty and are followed by a split at the bucket
    # of the smallest far distance, which the empty split measured
    off, col, w = _binary_tree((1 << 12) - 1, 100000, 5)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True) as g:
        h = host_csr(g)
        ref = oracle.sssp(h, 0)
        t0 = ctx.tuning
        try:
            for delta in (7, 1000, 50000):
                t = ctx.tuning
                t.sssp_delta = delta
                ctx.tuning = t
                d, st = grus.sssp(g, 0)
                print("delta %d rounds %d" % (delta, st.iterations), flush=True)
                assert (d == ref).all(), delta
            leaf = (1 << 12) - 2
            assert (grus.sssp(g, leaf)[0] == oracle.sssp(h, leaf)).all()
        finally:
            ctx.tuning = t0


@pytest.mark.parametrize("wmax", [1, 40])
def test_sssp_persistent_ring_spill(ctx, oracle, wmax):
    # a binary tree has degree <= 3 (the persistent path) and a frontier that
    # doubles per hop: the workgroup holding the source owes far more
    # vertices than its LDS ring holds, so owed vertices spill to the next
    # round's queue; distances still equal Dijkstra's, for every pass budget
    off, col, w = _binary_tree((1 << 17) - 1, wmax, 7)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True) as g:
        ref = oracle.sssp(host_csr(g), 0)
        t0 = ctx.tuning
        try:
            for hops in (0, 3):
                t = ctx.tuning
                t.sssp_local_hops = hops
                ctx.tuning = t
                d, st = grus.sssp(g, 0)
                print("wmax %d hops %d rounds %d" % (wmax, hops, st.iterations), flush=True)
                assert (d == ref).all()
        finally:
            ctx.tuning = t0
