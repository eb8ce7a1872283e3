"""Command-line tools and the autotuner on the GPU."""
import json
import os
import subprocess

import numpy as np
import pytest

import grus_framework_amd as grus

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "grus_framework_amd", "bin")


def load_dump(path, dtype):
    vals = np.loadtxt(path, dtype=np.float64)
    assert (vals[:, 0] == np.arange(len(vals))).all()
    return vals[:, 1].astype(dtype)


@pytest.mark.parametrize("app", ["bfs", "sssp", "cc", "pr"])
def test_cli_dump_matches_oracle(tmp_path, oracle, app):
    # a .gr input written by the product writer, read back by the CLI
    g = oracle.gen_rmat(11, weights=(1, 255), wseed=3)
    gr = str(tmp_path / "g.gr")
    grus.gr_write(gr, g.off, g.col, g.w)
    out = str(tmp_path / "out.txt")
    src = int(np.argmax(g.degrees()))
    r = subprocess.run([os.path.join(BIN, "grus_" + app), "--input", gr, "--symmetric",
                        "--src", str(src), "--output", out], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "itr " in r.stdout and "Done saving." in r.stdout
    if app == "pr":
        # the default dump has the reference's 6 significant digits
        # (include/print.cuh:36-45 streams floats at ostream's default
        # precision): each entry within half a unit of the 6th digit
        ref = oracle.pagerank(g)
        got = load_dump(out, np.float64)
        assert (np.abs(got - ref) <= 6e-6 * ref).all()
        assert all(len(line.split()[1].replace(".", "").lstrip("0").split("e")[0]) <= 6
                   for line in open(out))
        # --precise_output keeps 9 digits: the FP32 ranks within the L1 bound
        r = subprocess.run([os.path.join(BIN, "grus_pr"), "--input", gr, "--symmetric",
                            "--output", out, "--precise_output"], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        assert np.abs(load_dump(out, np.float64) - ref).sum() <= 1e-6
    else:
        ref = {"bfs": lambda: oracle.bfs(g, src), "sssp": lambda: oracle.sssp(g, src),
               "cc": lambda: oracle.cc(g)}[app]()
        assert (load_dump(out, np.uint64).astype(np.uint32) == ref).all()


def test_sweep_runs_each_dataset(tmp_path, oracle):
    # run_sweep.sh (the reference's run.sh) over two .w.gr files and one
    # missing name: one run per present file, each with its own source
    for name, scale in (("a", 9), ("b", 10)):
        g = oracle.gen_rmat(scale, weights=(1, 255), wseed=3)
        grus.gr_write(str(tmp_path / (name + ".w.gr")), g.off, g.col, g.w)
    r = subprocess.run(["bash", os.path.join(ROOT, "grus_framework_amd", "cli", "run_sweep.sh"),
                        "--data", str(tmp_path), "--app", "sssp", "a:3", "nope:0", "b:7",
                        "--", "--symmetric"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("itr ") == 2 and "== a (src 3)" in r.stdout
    assert "== b (src 7)" in r.stdout and "nope.w.gr not found" in r.stderr


def test_cli_placement_flags(tmp_path, oracle):
    out = str(tmp_path / "o.txt")
    for extra in (["--noopt", "--pf", "--ab"], ["--placement", "capped", "--hbm_cap_gb", "0.001"],
                  ["--cu_limit", "16"]):
        r = subprocess.run([os.path.join(BIN, "grus_bfs"), "--rmat", "12", "--src", "0",
                            "--output", out] + extra, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        g = oracle.gen_rmat(12)
        assert (load_dump(out, np.uint64).astype(np.uint32) == oracle.bfs(g, 0)).all()


def test_autotune_keeps_results_and_caches(ctx, tmp_path):
    cache = str(tmp_path / "tune.json")
    with grus.Graph.rmat(ctx, 16, weights=(1, 255), wseed=3) as g:
        s = [g.info.max_degree_vertex, 3, 77]
        before = [grus.bfs(g, x)[0] for x in s] + [grus.sssp(g, x)[0] for x in s]
        t0 = ctx.tuning
        try:
            tb = grus.autotune(g, "bfs", s, cache=cache)
            ts = grus.autotune(g, "sssp", s, cache=cache)
            assert tb.bfs_alpha in (2, 4, 8, 14, 24, 48, 96, 192)
            after = [grus.bfs(g, x)[0] for x in s] + [grus.sssp(g, x)[0] for x in s]
            for a, b in zip(before, after):
                assert (a == b).all()
            stored = json.load(open(cache))
            assert len(stored) == 2
            again = grus.autotune(g, "bfs", s, cache=cache)  # from the cache
            assert again.bfs_alpha == tb.bfs_alpha and again.sssp_delta == ts.sssp_delta
        finally:
            ctx.tuning = t0


def test_cu_mask_keeps_results(oracle):
    # a context restricted to a quarter of the CUs computes the same levels,
    # distances, labels and ranks; grids take the host-driven SSSP path
    # under a mask (a cooperative grid must be co-resident)
    with grus.Context(0) as c:
        active, total = c.cu_count
        assert active == total > 0
        with grus.Graph.rmat(c, 14, weights=(1, 255), wseed=3) as g, \
                grus.Graph.grid(c, 64, 80, weights=(1, 1000), wseed=2) as gg:
            s = g.info.max_degree_vertex
            full = [grus.bfs(g, s)[0], grus.sssp(g, s)[0], grus.cc(g)[0], grus.pagerank(g)[0],
                    grus.sssp(gg, 0)[0]]
            c.set_cu_limit(max(1, total // 4))
            assert c.cu_count == (max(1, total // 4), total)
            part = [grus.bfs(g, s)[0], grus.sssp(g, s)[0], grus.cc(g)[0], grus.pagerank(g)[0],
                    grus.sssp(gg, 0)[0]]
            for a, b in zip(full, part):
                assert (a.view(np.uint32) == b.view(np.uint32)).all()
            assert (part[4] == oracle.sssp(oracle.CSR(*gg.csr()), 0)).all()
            with pytest.raises(grus.GrusError) as e:
                c.set_cu_mask([0])
            assert e.value.code == grus.capi.E_INVALID
            c.set_cu_mask(None)
            assert c.cu_count == (total, total)
