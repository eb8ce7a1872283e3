"""GPU parity across placements: HBM, managed (with the reference's hints) and
capped (edge prefix in HBM, remainder in pinned host memory) must give the
same BFS levels, SSSP distances, CC labels and PageRank ranks."""
import numpy as np
import pytest

import grus_framework_amd as grus
from grus_framework_amd import capi

pytestmark = pytest.mark.gpu


def run_all(g, src):
    return (grus.bfs(g, src)[0], grus.bfs(g, src, "push")[0], grus.sssp(g, src)[0],
            grus.cc(g)[0], grus.pagerank(g)[0])


@pytest.fixture(scope="module")
def reference(ctx):
    g = grus.Graph.rmat(ctx, 15, weights=(1, 255), wseed=3)
    src = g.info.max_degree_vertex
    out = run_all(g, src)
    csr = g.csr()
    g.close()
    return src, csr, out


def check_same(got, ref):
    for a, b in zip(got[:4], ref[:4]):
        assert (a == b).all()
    assert np.abs(got[4] - ref[4]).sum() <= 1e-6


@pytest.mark.parametrize("frac", [0.0, 0.3, 0.9])
def test_capped_matches_hbm(ctx, reference, frac):
    src, (off, col, w), ref = reference
    V, E = len(off) - 1, len(col)
    # offsets + first neighbour per vertex + frac of the (col, weight) pairs
    cap = (V + 1) * 8 + V * 4 + int(frac * E) * 8
    pl = grus.place("capped", hbm_cap_bytes=cap)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True, placement=pl) as g:
        assert g.info.placement == grus.PLACE_CAPPED
        split = int(frac * E) // 64 * 64  # the HBM part is a whole number of 64-edge runs
        assert g.info.host_bytes == (E - split) * 8
        check_same(run_all(g, src), ref)
        o2, c2, w2 = g.csr()
        assert (c2 == col).all() and (w2 == w).all()


@pytest.mark.parametrize("flags", [dict(), dict(pf=True), dict(ab=True, rm=True),
                                   dict(pf=True, pl=True, ab=True, rm=True)])
def test_managed_matches_hbm(ctx, reference, flags):
    src, (off, col, w), ref = reference
    pl = grus.place("managed", **flags)
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True, placement=pl) as g:
        assert g.info.placement == grus.PLACE_MANAGED
        check_same(run_all(g, src), ref)


def test_capped_rmat_generator_and_errors(ctx, reference):
    src, _, ref = reference
    pl = grus.place("capped", hbm_cap_bytes=4 << 20)
    with grus.Graph.rmat(ctx, 15, weights=(1, 255), wseed=3, placement=pl) as g:
        assert g.info.host_bytes > 0
        check_same(run_all(g, src), ref)
    with pytest.raises(grus.GrusError) as e:
        grus.Graph.rmat(ctx, 15, placement=grus.place("capped", hbm_cap_bytes=1024))
    assert e.value.code == capi.E_BUDGET


def test_device_budget_caps_the_whole_footprint(ctx, reference):
    # a whole-device budget (the reference's tools/mem_limit.cu): the graph
    # takes what the budget leaves beside the app workspaces, every app runs
    # within it, and allocations past it fail with GRUS_E_BUDGET
    src, (off, col, w), ref = reference
    V, E = len(off) - 1, len(col)
    # a first budgeted placement releases the context's cached sort and scan
    # temporaries (held from earlier builds), so held0 counts what stays held
    with grus.Graph.from_csr(ctx, off, col, w, symmetric=True,
                             placement=grus.place("capped", device_budget_bytes=1 << 40)):
        pass
    ctx.set_device_budget(0)
    held0, _ = ctx.device_memory()
    # the documented reserve beside the graph (grus_c.h, device_budget_bytes)
    reserve = 48 * (V + 2) + 16 * (E // 2048 + 2) + (64 << 20)
    fixed = (V + 1) * 8 + V * 4  # offsets and first neighbours
    budget = held0 + reserve + fixed + (E // 2) * 8  # room for about half the edges
    try:
        pl = grus.place("capped", device_budget_bytes=budget)
        with grus.Graph.from_csr(ctx, off, col, w, symmetric=True, placement=pl) as g:
            info = g.info
            assert 0 < info.host_bytes < E * 8  # split between HBM and host
            held, _ = ctx.device_memory(reset_peak=True)
            assert held <= budget
            check_same(run_all(g, src), ref)
            _, peak = ctx.device_memory()
            assert peak <= budget
            # the budget is installed for the device: a graph whose build
            # needs more than the budget leaves fails before any device work
            with pytest.raises(grus.GrusError) as e:
                grus.Graph.rmat(ctx, 19)
            assert e.value.code == capi.E_BUDGET
        # a budget below what the library already holds is refused
        with pytest.raises(grus.GrusError) as e:
            ctx.set_device_budget(1024)
        assert e.value.code == capi.E_BUDGET
    finally:
        ctx.set_device_budget(0)
    # too small for the offsets and the workspaces: refused at placement
    with pytest.raises(grus.GrusError) as e:
        grus.Graph.from_csr(ctx, off, col, w, symmetric=True,
                            placement=grus.place("capped", device_budget_bytes=held0 + (1 << 20)))
    assert e.value.code == capi.E_BUDGET
    ctx.set_device_budget(0)
