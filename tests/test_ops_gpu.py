"""The public operator surface (include/grus_ops.hpp): BFS, SSSP and CC
written against it as user code (tests/ops_app/ops_apps.hip, in the shape
of the reference's samples) must give the oracle's levels, distances and
labels -- push over a work queue, pull over bitmaps, push over bitmaps from
a full frontier -- on RMAT, a grid, a directed graph and a capped graph."""
import ctypes as C
import os

import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
_APPS = None


def apps():
    global _APPS
    if _APPS is None:
        capi.lib()  # the engine first: the apps link against it
        path = os.path.join(HERE, "ops_app", "libops_apps.so")
        if not os.path.exists(path):
            raise RuntimeError("tests/ops_app/libops_apps.so is not built (run build())")
        L = C.CDLL(path)
        u32p = C.POINTER(C.c_uint32)
        L.ops_bfs.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, u32p]
        L.ops_sssp.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, u32p]
        L.ops_cc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, u32p]
        _APPS = L
    return _APPS


def run(name, ctx, g, *args):
    out = np.zeros(g.V, np.uint32)
    it = C.c_uint32()
    rc = getattr(apps(), name)(ctx.h, g.h, *args, out.ctypes.data_as(C.c_void_p), C.byref(it))
    capi.check(rc)
    return out, it.value


@pytest.mark.parametrize("scale", [10, 14])
def test_ops_apps_on_rmat(ctx, oracle, scale):
    with grus.Graph.rmat(ctx, scale, weights=(1, 255), wseed=3) as g:
        h = oracle.CSR(*g.csr())
        deg = h.degrees()
        srcs = [int(np.argmax(deg)), int(np.nonzero(deg)[0][-1])]
        for s in srcs:
            want = oracle.bfs(h, s)
            levels = int(want[want != grus.INFINIT].max()) + 1
            push, it = run("ops_bfs", ctx, g, s, 0)
            assert (push == want).all() and it == levels
            pull, it = run("ops_bfs", ctx, g, s, 1)
            assert (pull == want).all() and it == levels
            dist, _ = run("ops_sssp", ctx, g, s)
            assert (dist == oracle.sssp(h, s)).all()
        labels, _ = run("ops_cc", ctx, g)
        assert (labels == oracle.cc(h)).all()


def test_ops_apps_grid_directed_and_capped(ctx, oracle):
    with grus.Graph.grid(ctx, 60, 45, weights=(1, 1000), wseed=2) as g:
        h = oracle.CSR(*g.csr())
        assert (run("ops_bfs", ctx, g, 0, 0)[0] == oracle.bfs(h, 0)).all()
        assert (run("ops_bfs", ctx, g, 0, 1)[0] == oracle.bfs(h, 0)).all()
        assert (run("ops_sssp", ctx, g, 0)[0] == oracle.sssp(h, 0)).all()
    # directed: the pull reads the in-edges (CSC), the push the out-edges
    rng = np.random.default_rng(7)
    V = 3000
    pairs = np.unique(rng.integers(0, V, size=(20000, 2)).astype(np.int64), axis=0)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    off = np.zeros(V + 1, np.uint64)
    np.add.at(off, pairs[:, 0] + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    col = pairs[:, 1].astype(np.uint32)
    w = rng.integers(1, 100, len(col)).astype(np.uint32)
    h = oracle.CSR(off, col, w)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=False) as g:
        for s in (0, 17):
            assert (run("ops_bfs", ctx, g, s, 0)[0] == oracle.bfs(h, s)).all()
            assert (run("ops_bfs", ctx, g, s, 1)[0] == oracle.bfs(h, s)).all()
            assert (run("ops_sssp", ctx, g, s)[0] == oracle.sssp(h, s)).all()
    # capped: most edges read from host memory through graph_t::dst / weight
    hc = oracle.gen_rmat(12, weights=(1, 255), wseed=3)
    cap = (hc.V + 1) * 8 + hc.V * 4 + (hc.E // 4) * 8
    with grus.Graph.from_csr(ctx, hc.off, hc.col, hc.w, symmetric=True,
                             placement=grus.place("capped", hbm_cap_bytes=cap)) as g:
        assert g.info.host_bytes > 0
        s = int(np.argmax(hc.degrees()))
        assert (run("ops_bfs", ctx, g, s, 0)[0] == oracle.bfs(hc, s)).all()
        assert (run("ops_bfs", ctx, g, s, 1)[0] == oracle.bfs(hc, s)).all()
        assert (run("ops_sssp", ctx, g, s)[0] == oracle.sssp(hc, s)).all()
        assert (run("ops_cc", ctx, g)[0] == oracle.cc(hc)).all()


def test_ops_errors(ctx):
    with grus.Graph.rmat(ctx, 8) as g:
        out = np.zeros(g.V, np.uint32)
        it = C.c_uint32()
        rc = apps().ops_bfs(ctx.h, g.h, g.V, 0, out.ctypes.data_as(C.c_void_p), C.byref(it))
        assert rc == capi.E_INVALID
