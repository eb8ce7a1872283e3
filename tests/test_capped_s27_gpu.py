"""Capped-HBM placement at its benchmark size: RMAT scale 27 (edge factor 16,
a/b/c = .57/.19/.19, seed 1, symmetrized: 134,217,728 vertices, ~4.2 G
directed edges, ~17 GB of column ids) with the whole device capped at
8 GiB, so most edges are read from pinned host memory over the host link.

- The resident run (everything in HBM) is checked against the oracle: BFS
  levels from the max-degree vertex by the O(E) BFS certificate (the levels
  are the hop distances), PageRank after two iterations within L1 1e-6 of
  the FP64 oracle.
- The capped run must give the same levels bit for bit and ranks within
  L1 1e-6, while the bytes the library holds on the device never exceed the
  cap (graph part, app workspaces, everything).
"""
import numpy as np
import pytest

import grus_framework_amd as grus

INF = grus.INFINIT
CAP = 8 << 30
PR_TOL = 1e-6


@pytest.mark.gpu
@pytest.mark.timeout(1200)
def test_rmat_s27_capped_at_8gib_matches_resident(ctx, oracle):
    t_default = ctx.tuning
    print("s27 resident: generate", flush=True)
    with grus.Graph.rmat(ctx, 27) as g:
        src = g.info.max_degree_vertex
        print("s27 resident: bfs from", src, flush=True)
        lv_res, st_res = grus.bfs(g, src)
        # the plain pull, the layout the capped placement runs
        t = ctx.tuning
        t.pr_layout = 2
        ctx.tuning = t
        try:
            print("s27 resident: pagerank, 2 iterations", flush=True)
            pr_res, _ = grus.pagerank(g, 0.85, 2)
        finally:
            ctx.tuning = t_default
        print("s27: copy the CSR to the host", flush=True)
        off, col, _ = g.csr()
    h = oracle.CSR(off, col)
    del off, col
    print("s27: bfs certificate", flush=True)
    assert oracle.bfs_certify(h, src, lv_res) == (0, 0)
    reached = lv_res != INF
    assert st_res.vertices_reached == int(reached.sum())
    assert st_res.edges_reached == int(h.degrees()[reached].sum())
    print("s27: pagerank oracle, 2 iterations", flush=True)
    assert np.abs(pr_res.astype(np.float64) - oracle.pagerank(h, 0.85, 2)).sum() <= PR_TOL
    E = h.E
    del h

    held0, _ = ctx.device_memory()
    try:
        print("s27 capped at 8 GiB: generate and place", flush=True)
        pl = grus.place("capped", device_budget_bytes=CAP)
        with grus.Graph.rmat(ctx, 27, placement=pl) as gc:
            info = gc.info
            assert info.placement == grus.PLACE_CAPPED and info.E == E
            assert info.host_bytes > E * 4 // 2  # most edges stream from the host
            held, _ = ctx.device_memory(reset_peak=True)
            assert held <= CAP
            print("s27 capped: bfs (graph %.2f GB on the device, %.2f GB of edges on the host)"
                  % (info.device_bytes / 1e9, info.host_bytes / 1e9), flush=True)
            lv_cap, st = grus.bfs(gc, src)
            assert (lv_cap == lv_res).all()
            assert st.edges_reached == st_res.edges_reached
            print("s27 capped: pagerank, 2 iterations", flush=True)
            pr_cap, _ = grus.pagerank(gc, 0.85, 2)
            assert np.abs(pr_cap.astype(np.float64) - pr_res).sum() <= PR_TOL
            _, peak = ctx.device_memory()
            print("s27 capped: peak device bytes %d of %d" % (peak, CAP), flush=True)
            assert peak <= CAP
    finally:
        ctx.set_device_budget(0)
