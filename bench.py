"""Benchmark: BFS GTEPS on RMAT scale 24 (one MI355X), plus SSSP / PageRank / CC.

python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step" = one BFS traversal from one source): RMAT scale 24,
edge factor 16, a/b/c = .57/.19/.19, seed 1, symmetrized, deduplicated,
self-loops dropped (V = 16,777,216, E = 520,761,196 directed edges), built
and resident in HBM before timing.  Sources: the max-degree vertex and 16
seeded random non-isolated vertices, cycled.  GTEPS = sum of E_r (directed
edges of the reached component) over the timed steps / wall time.

Prints one JSON line.  The roofline object covers the BFS traversal; the CPU
baseline is the oracle's OpenMP direction-optimizing BFS on this host's
cores over a bounded sample of the same sources.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pick_sources(deg, n, seed):
    rng = np.random.default_rng(seed)
    nz = np.nonzero(deg)[0]
    return [int(np.argmax(deg))] + [int(x) for x in rng.choice(nz, n, replace=False)]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=34)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU baseline sample")
    ap.add_argument("--no-extra", action="store_true", help="skip SSSP / PageRank / CC")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path "
              "(every box here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import grus_framework_amd as grus
    from oracle import oracle_py as O

    ctx = grus.Context(0)
    t0 = time.perf_counter()
    g = grus.Graph.rmat(ctx, args.scale)
    log("rmat s%d built on device in %.2fs: V=%d E=%d" % (args.scale, time.perf_counter() - t0,
                                                          g.V, g.E))
    off, col, _ = g.csr()
    h = O.CSR(off, col)
    deg = h.degrees()
    srcs = pick_sources(deg, 16, seed=5)

    for i in range(args.warmup):
        grus.bfs(g, srcs[i % len(srcs)], copy=False)
    er = vr = 0
    dev_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = grus.bfs(g, srcs[i % len(srcs)], copy=False)  # returns when the stream is idle
        er += st.edges_reached
        vr += st.vertices_reached
        dev_ms += st.device_ms
    wall = time.perf_counter() - t0
    gteps = er / wall / 1e9
    algo_bytes = 4 * er + 16 * vr + 4 * g.V * args.steps
    achieved = algo_bytes / (dev_ms * 1e-3) / 1e9
    log("bfs: %d steps %.3f ms/step, %.1f GTEPS (device %.3f ms/step)" %
        (args.steps, wall / args.steps * 1e3, gteps, dev_ms / args.steps))

    # CPU baseline: OpenMP direction-optimizing BFS on a bounded sample
    threads = O.max_threads()
    cpu_er, cpu_ms, n_cpu = 0, 0.0, 0
    for s in srcs:
        lv, ms = O.bl_bfs(h, s, threads)
        cpu_er += int(deg[lv != O.INFINIT].sum())
        cpu_ms += ms
        n_cpu += 1
        if cpu_ms > args.cpu_seconds * 1e3:
            break
    cpu_gteps = cpu_er / (cpu_ms * 1e-3) / 1e9
    log("cpu bfs: %d sources %.1f ms each, %.2f GTEPS on %d threads (%s)" %
        (n_cpu, cpu_ms / n_cpu, cpu_gteps, threads, cpu_model()))

    extra = {}
    if not args.no_extra:
        extra = run_extra(grus, O, ctx, g, h, threads, args)
    g.close()

    line = {
        "metric": "BFS GTEPS at RMAT s24 (directed edges of the reached component / s)",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (RMAT generated on the device, seed 1)",
        "config": {"workload": "bfs_rmat_s%d_ef16" % args.scale, "V": int(g.V),
                   "E": int(h.E), "sources": len(srcs), "direction": "auto",
                   "graph500_gteps": round(gteps / 2, 3)},
        "roofline": {"kernel": "bfs traversal (all launches of one source)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_step": int(algo_bytes // args.steps)},
        "cpu_baseline": {"value": round(cpu_gteps, 3), "unit": "GTEPS", "cores": threads,
                         "kind": "port",
                         "sample": "OpenMP direction-optimizing BFS from %d of the %d sources "
                                   "on the full graph (%s)" % (n_cpu, len(srcs), cpu_model())},
        "gpu_vs_cpu": round(gteps / cpu_gteps, 2) if cpu_gteps else None,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


def run_extra(grus, O, ctx, g, h, threads, args):
    out = {}
    # PageRank pull, 20 iterations, FP32 (same RMAT graph)
    grus.pagerank(g, copy=False)
    reps = 3
    ms = 0.0
    for _ in range(reps):
        _, st = grus.pagerank(g, copy=False)
        ms += st.device_ms
    ms /= reps
    B = 20 * (8 * (g.V + 1) + 8 * g.E + 12 * g.V)
    _, cms = O.bl_pagerank(h, 0.85, 2, threads)  # 2-iteration CPU sample, scaled
    cpu_gbs = (B / 10) / (cms * 1e-3) / 1e9
    out["pagerank"] = {"ms_20_iters": round(ms, 3), "GBps_algorithmic": round(B / ms / 1e6, 1),
                       "frac_hbm_peak": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
                       "cpu_GBps": round(cpu_gbs, 2), "cpu_sample": "2 iterations, scaled"}
    log("pagerank: %.3f ms (%.1f GB/s)" % (ms, B / ms / 1e6))
    # CC (hooking) on the same graph
    grus.cc(g, copy=False)
    ms = min(grus.cc(g, copy=False)[1].device_ms for _ in range(3))
    _, cms = O.bl_cc(h, threads)
    out["cc"] = {"ms": round(ms, 3), "GTEPS": round(g.E / ms / 1e6, 2), "cpu_ms": round(cms, 1)}
    log("cc: %.3f ms (cpu %.1f ms)" % (ms, cms))
    # SSSP on the 4096x4096 weighted grid (source 0)
    with grus.Graph.grid(ctx, 4096, 4096, weights=(1, 1000), wseed=2) as gg:
        _, st = grus.sssp(gg, 0, copy=False)
        ms = st.device_ms
        gteps = st.edges_reached / ms / 1e6
        hh = O.CSR(*gg.csr())
        _, cms = O.bl_sssp(hh, 0, 4000, threads)
        cpu_gteps = st.edges_reached / cms / 1e6
    out["sssp_grid4096"] = {"ms": round(ms, 2), "GTEPS": round(gteps, 3), "rounds": st.iterations,
                            "cpu_ms": round(cms, 1), "cpu_GTEPS": round(cpu_gteps, 3)}
    log("sssp grid: %.2f ms (cpu %.1f ms)" % (ms, cms))
    return out


if __name__ == "__main__":
    main()
