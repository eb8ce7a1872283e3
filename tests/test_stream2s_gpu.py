"""Two-stage active-subgraph streaming BFS (grus_bfs_2s; the reference's
samples/bfs_2s with Frontier_2S / Subgraph): the frontier's rows are
gathered into a device staging buffer, in pieces when they do not fit, and
expanded from there.  Levels must equal the oracle's for every staging size
(pieces that cut rows in two), every placement (rows read from HBM or from
pinned host memory) and directed graphs."""
import os
import subprocess

import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("stage", [0, 777, 65536])
def test_bfs_2s_rmat_stage_sizes(ctx, oracle, stage):
    with grus.Graph.rmat(ctx, 14) as g:
        h = oracle.CSR(*g.csr())
        deg = h.degrees()
        for s in (int(np.argmax(deg)), int(np.nonzero(deg)[0][-1]), int(np.nonzero(deg == 0)[0][0])):
            lv, st = grus.bfs_2s(g, s, stage)
            want = oracle.bfs(h, s)
            assert (lv == want).all(), (s, stage)
            reached = want != grus.INFINIT
            assert st.vertices_reached == reached.sum()
            assert st.edges_reached == deg[reached].sum()
            assert st.iterations == int(want[reached].max()) + 1  # levels run
            if stage:
                assert st.reserved[0] == min(stage, g.E)


def test_bfs_2s_capped_grid_and_directed(ctx, oracle):
    hc = oracle.gen_rmat(15)
    cap = (hc.V + 1) * 8 + hc.V * 4 + (hc.E // 5) * 4  # a fifth of the edges in HBM
    pl = grus.place("capped", hbm_cap_bytes=cap)
    with grus.Graph.from_csr(ctx, hc.off, hc.col, symmetric=True, placement=pl) as g:
        assert g.info.host_bytes > hc.E * 2
        s = int(np.argmax(hc.degrees()))
        want = oracle.bfs(hc, s)
        for stage in (0, 4096):
            assert (grus.bfs_2s(g, s, stage)[0] == want).all()
        assert (grus.bfs(g, s, "push")[0] == want).all()
    with grus.Graph.grid(ctx, 70, 50) as g:
        rr, cc = np.divmod(np.arange(70 * 50), 50)
        assert (grus.bfs_2s(g, 0, 300)[0] == rr + cc).all()
    rng = np.random.default_rng(4)
    V = 2000
    key = np.unique(rng.integers(0, V, 12000).astype(np.uint64) << np.uint64(32) |
                    rng.integers(0, V, 12000).astype(np.uint64))
    src = (key >> np.uint64(32)).astype(np.int64)
    col = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, src + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    hd = oracle.CSR(off, col)
    with grus.Graph.from_csr(ctx, off, col, symmetric=False) as g:
        for s in (0, 99, 1999):
            assert (grus.bfs_2s(g, s, 500)[0] == oracle.bfs(hd, s)).all()
        with pytest.raises(grus.GrusError) as e:
            grus.bfs_2s(g, V)
        assert e.value.code == capi.E_INVALID


def test_bfs_2s_cli(tmp_path, oracle):
    out = str(tmp_path / "lv.txt")
    r = subprocess.run([os.path.join(ROOT, "grus_framework_amd", "bin", "grus_bfs_2s"), "--rmat",
                        "12", "--src", "5", "--stage_edges", "3000", "--output", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "itr " in r.stdout
    vals = np.loadtxt(out, dtype=np.uint64)
    assert (vals[:, 1].astype(np.uint32) == oracle.bfs(oracle.gen_rmat(12), 5)).all()
