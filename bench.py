"""Benchmark of the traversal engine on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extra] [--no-cpu]

Headline workload (one step = one BFS traversal from one source): RMAT scale
24, edge factor 16, a/b/c = .57/.19/.19, seed 1, symmetrized, deduplicated,
self-loops dropped (V = 16,777,216, E = 520,761,196 directed edges), built
on the device and resident in HBM before timing.  Sources: the max-degree
vertex and 16 seeded random non-isolated vertices, cycled; the direction
thresholds are the library defaults (alpha 14, beta 64, 32 bottom-up edges
per lane: what the autotuner picks on this graph in most runs, pinned so
the measured HBM traffic in profiles/traffic.json belongs to the same
schedule; --autotune tunes them first on a different seeded sample).  value =
GTEPS = sum over the timed steps of E_r (directed edges of the reached
component) / wall time of the timed loop (each call returns when the device
is idle).

Also reported (the rest of the metric): SSSP GTEPS on the 4096x4096 grid
and on weighted RMAT s24, PageRank GB/s and CC on RMAT s24, batched
serving on RMAT s22 (queries/s, p50/p99, vs a one-query-at-a-time loop) and
capped-HBM placement on RMAT s27 -- each next to an OpenMP CPU baseline or
the resident run where one applies.  One JSON line on stdout.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


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


class Regions:
    """Host-clock windows around the timed loops.

    Each window records CLOCK_MONOTONIC and CLOCK_BOOTTIME in ns at its
    begin and end; `--regions-out FILE` writes them as JSON.  rocprofv3
    stamps every dispatch on the host clock, so
    `python -m grus_framework_amd.rocprof_summary` can keep exactly the
    dispatches of the timed loops (not the graph build, the autotuner's
    trial runs or the warm-up) when it summarises a kernel trace or a
    counter run of this script.
    """

    def __init__(self):
        self.windows = []

    @staticmethod
    def _now():
        return {"monotonic": time.clock_gettime_ns(time.CLOCK_MONOTONIC),
                "boottime": time.clock_gettime_ns(time.CLOCK_BOOTTIME)}

    def begin(self, name):
        self.windows.append({"name": name, "begin": self._now()})

    def end(self, units, config=None):
        self.windows[-1]["end"] = self._now()
        self.windows[-1]["units"] = units
        if config is not None:  # settings the window ran with (kept with its traffic)
            self.windows[-1]["config"] = config

    def write(self, path):
        if path:
            with open(path, "w") as f:
                json.dump(self.windows, f, indent=1)


REGIONS = Regions()


class Watchdog:
    """Guards the extra legs after the headline is measured.

    Each leg announces itself (`leg`), which logs its name.  A leg that has
    not finished after `limit_s` seconds (a stuck device call cannot be
    interrupted from Python) is abandoned: the JSON line is printed with
    that leg marked as having no result, and the process exits, so the
    headline and the legs that finished are still reported."""

    def __init__(self, limit_s):
        self.limit_s = limit_s
        self.lock = threading.Lock()
        self.printed = False
        self.line = None
        self.name = None
        self.t0 = time.time()
        if limit_s > 0:
            threading.Thread(target=self._watch, daemon=True).start()

    def leg(self, name):
        self.name, self.t0 = name, time.time()
        log("leg %s" % name)

    def emit(self, line):
        with self.lock:
            if not self.printed:
                self.printed = True
                print(json.dumps(line), flush=True)

    def _watch(self):
        while True:
            time.sleep(2)
            line, name = self.line, self.name
            if line is None or name is None or time.time() - self.t0 < self.limit_s:
                continue
            # the main thread is stuck inside the leg, so these copies do not
            # race with it
            line = dict(line)
            line["extra"] = dict(line["extra"])
            line["extra"][name] = {"error": "no result within %d s; leg abandoned" %
                                            self.limit_s}
            log("leg %s: no result within %d s, reporting without it" % (name, self.limit_s))
            finish_line(line)
            self.emit(line)
            REGIONS.write(ARGS_REGIONS_OUT[0])
            os._exit(0)


def finish_line(line):
    extra = line["extra"]
    if "pagerank" in extra and "roofline" in extra["pagerank"]:
        line["roofline_pagerank"] = extra["pagerank"].pop("roofline")


ARGS_REGIONS_OUT = [""]


def copy_calibration():
    """One plain device-to-device copy of exactly 1 GiB inside a profiled
    region: FETCH_SIZE / WRITE_SIZE of this dispatch give the factors that
    turn the counters into bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(1 << 30)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for rc in (hip.hipMalloc(ctypes.byref(a), n), hip.hipMalloc(ctypes.byref(b), n),
               hip.hipMemset(a, 1, n), hip.hipDeviceSynchronize()):
        if rc != 0:
            raise RuntimeError("copy calibration: hip error %d" % rc)
    REGIONS.begin("copy_1GiB")
    rc = hip.hipMemcpy(b, a, n, 3)  # 3 = hipMemcpyDeviceToDevice
    rc = rc or hip.hipDeviceSynchronize()
    REGIONS.end(1)
    hip.hipFree(a)
    hip.hipFree(b)
    if rc != 0:
        raise RuntimeError("copy calibration: hip error %d" % rc)
    log("copy calibration: 1 GiB device-to-device copy done")


def traffic(key, config=None):
    """Measured HBM bytes per unit of a timed window (rocprofv3 FETCH_SIZE and
    WRITE_SIZE, scaled), or None when not recorded or when it was recorded
    with other settings than `config` (the window's settings in this run)."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)[key]
        if config is not None and t.get("config") != config:
            return None
        return int(t["fetch_bytes"] + t["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=34)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="budget of the CPU baseline sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only (no SSSP / PageRank / CC / serving / capped)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the library's default thresholds (the default)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the BFS thresholds on another sample of sources first")
    ap.add_argument("--bfs-tuning", default="",
                    help="alpha,beta,cutoff: use these BFS thresholds instead of autotuning "
                         "(e.g. the values an unprofiled run's autotuner printed)")
    ap.add_argument("--only", default="",
                    help="comma list of extras to run: pagerank,cc,sssp (or sssp_rmat, "
                         "sssp_grid),cu,serving,capped; batched (the batched engines "
                         "alone, run only when named)")
    ap.add_argument("--leg-limit", type=float, default=240.0,
                    help="seconds after which an extra leg with no result is abandoned "
                         "(0 = never)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB device-to-device copy in its own timed window")
    ap.add_argument("--regions-out", default="",
                    help="write the host-clock windows of the timed loops to this JSON file")
    ap.add_argument("--leg-child", default="", choices=["", "serving", "cu", "capped"],
                    help="run one extra leg in this process and print its result as JSON "
                         "(what the full run starts as a child process; also the target "
                         "for profiling a leg)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path "
              "(every box here has one GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import grus_framework_amd as grus
    from oracle import oracle_py as O

    if args.leg_child:
        if args.leg_child == "serving":
            res = run_serving(grus)
        else:
            cctx = grus.Context(0)
            if args.leg_child == "cu":
                with grus.Graph.rmat(cctx, args.scale) as cg:
                    res = run_cu_scaling(grus, cctx, cg, args)
            else:
                try:
                    res = run_capped(grus, cctx)
                except grus.GrusError as e:  # e.g. not enough device or host memory
                    res = {"error": str(e)}
            cctx.close()
        REGIONS.write(args.regions_out)
        print(json.dumps(res), flush=True)
        return

    if args.copy_calibration:
        copy_calibration()
    ctx = grus.Context(0)
    t0 = time.perf_counter()
    g = grus.Graph.rmat(ctx, args.scale)
    log("rmat s%d built on device in %.2fs: V=%d E=%d" % (args.scale, time.perf_counter() - t0,
                                                          g.V, g.E))
    off, col, _ = g.csr()
    h = O.CSR(off, col)
    deg = h.degrees()
    srcs = pick_sources(deg, 16, seed=5)
    tuned = None
    if args.bfs_tuning:
        a, b, k = args.bfs_tuning.split(",")
        t = ctx.tuning
        t.bfs_alpha, t.bfs_beta, t.bu_thread_edges = float(a), float(b), int(k)
        ctx.tuning = t
        tuned = {"bfs_alpha": t.bfs_alpha, "bfs_beta": t.bfs_beta,
                 "bu_thread_edges": t.bu_thread_edges, "source": "--bfs-tuning"}
    elif args.autotune and not args.no_tune:
        tune_srcs = pick_sources(deg, 8, seed=99)[1:]
        t = grus.autotune(g, "bfs", tune_srcs)
        tuned = {"bfs_alpha": t.bfs_alpha, "bfs_beta": t.bfs_beta,
                 "bu_thread_edges": t.bu_thread_edges}
        log("autotuned on other sources: %s" % tuned)

    for i in range(args.warmup):
        grus.bfs(g, srcs[i % len(srcs)], copy=False)
    er = vr = 0
    dev_ms = 0.0
    REGIONS.begin("bfs_rmat_s%d_traversal" % args.scale)
    t0 = time.perf_counter()
    for i in range(args.steps):
        _, st = grus.bfs(g, srcs[i % len(srcs)], copy=False)  # returns when the stream is idle
        er += st.edges_reached
        vr += st.vertices_reached
        dev_ms += st.device_ms
    wall = time.perf_counter() - t0
    tc = ctx.tuning
    bfs_config = {"bfs_tuning": [tc.bfs_alpha, tc.bfs_beta, tc.bu_thread_edges]}
    REGIONS.end(args.steps, bfs_config)
    gteps = er / wall / 1e9
    algo_bytes = (4 * er + 16 * vr + 4 * g.V * args.steps) / args.steps
    achieved = algo_bytes / (dev_ms / args.steps * 1e-3) / 1e9
    log("bfs: %d steps %.3f ms/step, %.1f GTEPS (device %.3f ms/step)" %
        (args.steps, wall / args.steps * 1e3, gteps, dev_ms / args.steps))

    threads = O.max_threads()
    cpu = None
    if not args.no_cpu:
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
        cpu = {"value": round(cpu_gteps, 3), "unit": "GTEPS", "cores": threads, "kind": "port",
               "sample": "OpenMP direction-optimizing BFS (oracle/baseline_omp.c) from %d of "
                         "the %d sources on the full graph, %s" % (n_cpu, len(srcs), cpu_model())}

    extra = {}
    bfs_traffic = traffic("bfs_rmat_s%d_traversal" % args.scale, bfs_config)
    line = {
        "metric": "BFS GTEPS at RMAT s24 (directed edges of the reached component / s); "
                  "SSSP GTEPS, PageRank GB/s, serving queries/s in extra",
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
        "config": {"workload": "bfs_rmat_s%d_ef16" % args.scale, "V": int(g.V), "E": int(h.E),
                   "sources": len(srcs),
                   "direction": ("auto (" + ("pinned" if args.bfs_tuning else "autotuned") + ")")
                   if tuned else "auto (library defaults)",
                   "tuning": tuned, "graph500_gteps": round(gteps / 2, 3)},
        "roofline": {"kernel": "BFS traversal of one source (all its launches; k_pull "
                               "dominates)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": bfs_traffic,
                     "traffic_config": bfs_config,
                     # HBM bytes the counters saw per traversal / its device time
                     "frac_counters": round(bfs_traffic / (dev_ms / args.steps * 1e-3) / 1e9
                                            / HBM_PEAK_GBS, 4) if bfs_traffic else None,
                     "algorithmic_bytes": int(algo_bytes),
                     "device_ms": round(dev_ms / args.steps, 4)},
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(gteps / cpu["value"], 2) if cpu else None,
        "extra": extra,
    }
    ARGS_REGIONS_OUT[0] = args.regions_out
    watch = Watchdog(args.leg_limit if not args.no_extra else 0)
    watch.line = line
    if not args.no_extra:
        run_extra(grus, O, ctx, g, h, threads, args, extra, watch)
    g.close()
    # the capped leg caps the whole device: it runs with no other graph held
    only = set(x for x in args.only.split(",") if x)
    if not args.no_extra and (not only or "capped" in only):
        watch.leg("capped_rmat_s27")
        extra["capped_rmat_s27"] = run_leg_child("capped", args)
    watch.name = None  # every leg done
    finish_line(line)
    watch.emit(line)
    REGIONS.write(args.regions_out)
    ctx.close()


def run_extra(grus, O, ctx, g, h, threads, args, out, watch):
    only = set(x for x in args.only.split(",") if x)

    def want(name):
        return not only or name in only

    if want("pagerank"):
        watch.leg("pagerank")
        run_pagerank(grus, O, g, h, threads, args, out)
    if want("cc"):
        watch.leg("cc")
        run_cc(grus, O, g, h, threads, args, out)
    if want("sssp") or want("sssp_rmat") or want("sssp_grid"):
        watch.leg("sssp")
        run_sssp(grus, O, ctx, threads, args, out,
                 rmat=want("sssp") or want("sssp_rmat"), grid=want("sssp") or want("sssp_grid"))
    if want("serving"):
        watch.leg("serving_rmat_s22")
        out["serving_rmat_s22"] = run_leg_child("serving", args)
    if want("cu"):
        watch.leg("bfs_cu_scaling")
        out["bfs_cu_scaling"] = run_leg_child("cu", args)
    if "batched" in only:  # named only: a kernel-trace target for the batched engines
        watch.leg("batched_rmat_s22")
        out["batched_rmat_s22"] = run_batched(grus)


def run_batched(grus):
    """The batched engines alone, without the server: three 128-query SSSP
    batches and three 64-query BFS batches on the serving graph (RMAT s22,
    w in [1,255], sources as in the serving leg), each set in its own timed
    window, after one untimed batch of each."""
    ctx = grus.Context(0)
    g = grus.Graph.rmat(ctx, 22, weights=(1, 255), wseed=3)
    off, _, _ = g.csr()
    nz = np.nonzero(np.diff(off.astype(np.int64)))[0]
    srcs = [int(x) for x in np.random.default_rng(5).choice(nz, 3 * 128 + 3 * 64)]
    res = {}
    grus.sssp_multi(g, srcs[:128], copy=False)
    REGIONS.begin("batched_sssp128_rmat_s22_batch")
    ms = [grus.sssp_multi(g, srcs[128 * i:128 * (i + 1)], copy=False)[1].device_ms
          for i in range(3)]
    REGIONS.end(3)
    res["sssp_batch128_ms"] = round(float(np.median(ms)), 2)
    b0 = 3 * 128
    grus.bfs_multi(g, srcs[b0:b0 + 64], copy=False)
    REGIONS.begin("batched_bfs64_rmat_s22_batch")
    ms = [grus.bfs_multi(g, srcs[b0 + 64 * i:b0 + 64 * (i + 1)], copy=False)[1].device_ms
          for i in range(3)]
    REGIONS.end(3)
    res["bfs_batch64_ms"] = round(float(np.median(ms)), 2)
    log("batched s22: SSSP 128 queries %.1f ms, BFS 64 queries %.1f ms" %
        (res["sssp_batch128_ms"], res["bfs_batch64_ms"]))
    g.close()
    ctx.close()
    return res


def run_pagerank(grus, O, g, h, threads, args, out):
    """PageRank, 20 iterations, FP32 ranks, on the headline RMAT graph (the
    library's automatic layout: binned push on HBM-resident graphs)."""
    grus.pagerank(g, copy=False)
    reps = 3
    ms = 0.0
    REGIONS.begin("pagerank_rmat_s%d_iteration" % args.scale)
    for _ in range(reps):
        _, st = grus.pagerank(g, copy=False)
        ms += st.device_ms
    REGIONS.end(reps * 20)
    ms /= reps
    per_iter = 8 * (g.V + 1) + 8 * g.E + 12 * g.V
    B = 20 * per_iter
    pr = {"ms_20_iters": round(ms, 3), "GBps_algorithmic": round(B / ms / 1e6, 1),
          "frac_hbm_peak": round(B / ms / 1e6 / HBM_PEAK_GBS, 4)}
    pr["roofline"] = {"kernel": "one PageRank iteration, binned push layout "
                                "(pr::k_bp_scatter + pr::k_bp_accum + pr::k_bp_finish)",
                      "bound": "hbm",
                      "achieved": round(per_iter / (ms / 20 * 1e-3) / 1e9, 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(per_iter / (ms / 20 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": traffic("pagerank_rmat_s24_iteration"),
                      "algorithmic_bytes": int(per_iter),
                      "note": "device time of the 20-iteration call / 20 (one init "
                              "launch included); the unit is the iteration"}
    if not args.no_cpu:
        _, cms = O.bl_pagerank(h, 0.85, 2, threads)  # 2-iteration CPU sample, scaled
        pr["cpu_GBps"] = round((B / 10) / (cms * 1e-3) / 1e9, 2)
        pr["cpu_sample"] = "2 of 20 iterations, %d threads" % threads
    out["pagerank"] = pr
    log("pagerank: %.3f ms (%.1f GB/s)" % (ms, B / ms / 1e6))


def run_cc(grus, O, g, h, threads, args, out):
    """CC (hooking) on the headline RMAT graph."""
    grus.cc(g, copy=False)
    ms = min(grus.cc(g, copy=False)[1].device_ms for _ in range(3))
    cc = {"ms": round(ms, 3), "GTEPS": round(g.E / ms / 1e6, 2)}
    if not args.no_cpu:
        _, cms = O.bl_cc(h, threads)
        cc["cpu_ms"] = round(cms, 1)
    out["cc"] = cc
    log("cc: %.3f ms" % ms)


def run_sssp(grus, O, ctx, threads, args, out, rmat=True, grid=True):
    """SSSP on weighted RMAT s24 (w in [1,255]) and on the 4096x4096 grid."""
    if rmat:
        run_sssp_rmat(grus, O, ctx, threads, args, out)
    if grid:
        run_sssp_grid(grus, O, ctx, threads, args, out)


def run_sssp_rmat(grus, O, ctx, threads, args, out):
    with grus.Graph.rmat(ctx, args.scale, weights=(1, 255), wseed=3) as gw:
        s = gw.info.max_degree_vertex
        grus.sssp(gw, s, copy=False)  # first call builds the binned layout
        ms = []
        REGIONS.begin("sssp_rmat_s%d_traversal" % args.scale)
        for _ in range(3):
            dist, st = grus.sssp(gw, s, copy=True)
            ms.append(st.device_ms)
        REGIONS.end(3)
        ms = float(np.median(ms))
        rm = {"ms": round(ms, 2), "GTEPS": round(st.edges_reached / ms / 1e6, 3),
              "rounds": st.iterations, "pull_rounds": st.pull_iterations,
              "delta": int(st.reserved[1])}
        if not args.no_cpu:
            hw = O.CSR(*gw.csr())
            cdist, cms = O.bl_sssp(hw, s, int(st.reserved[1]), threads)
            rm["cpu_ms"] = round(cms, 1)
            rm["cpu_GTEPS"] = round(st.edges_reached / cms / 1e6, 3)
            rm["cpu_threads"] = threads
            rm["distances_equal_cpu"] = bool((cdist == dist).all())
            del hw, cdist
        out["sssp_rmat_s%d" % args.scale] = rm
        log("sssp rmat: %.2f ms%s" % (ms, (", cpu %.1f ms (equal: %s)" % (
            rm["cpu_ms"], rm["distances_equal_cpu"])) if "cpu_ms" in rm else ""))


def run_sssp_grid(grus, O, ctx, threads, args, out):
    """SSSP on the 4096x4096 grid, weights in [1,1000], source 0."""
    with grus.Graph.grid(ctx, 4096, 4096, weights=(1, 1000), wseed=2) as gg:
        grus.sssp(gg, 0, copy=False)
        REGIONS.begin("sssp_grid4096_traversal")
        _, st = grus.sssp(gg, 0, copy=False)
        REGIONS.end(1)
        ms = st.device_ms
        grid = {"ms": round(ms, 2), "GTEPS": round(st.edges_reached / ms / 1e6, 3),
                "rounds": st.iterations}
        if not args.no_cpu:
            hh = O.CSR(*gg.csr())
            _, cms = O.bl_sssp(hh, 0, 4000, threads)
            grid["cpu_ms"] = round(cms, 1)
            grid["cpu_GTEPS"] = round(st.edges_reached / cms / 1e6, 3)
        out["sssp_grid4096"] = grid
        log("sssp grid: %.2f ms" % ms)


def run_cu_scaling(grus, ctx, g, args):
    """BFS GTEPS on the headline graph with the context restricted to the
    first 1/8, 1/4, 1/2 and all of the CUs (co-location sizing; the analogue
    of the reference's SM-limit experiments)."""
    off = g.csr()[0]
    deg = np.diff(off.astype(np.int64))
    srcs = pick_sources(deg, 16, seed=5)
    total = ctx.cu_count[1]
    res = {}
    try:
        for frac in (8, 4, 2, 1):
            n = max(1, total // frac)
            ctx.set_cu_limit(n)
            for s in srcs[:2]:
                grus.bfs(g, s, copy=False)
            er, t0 = 0, time.perf_counter()
            for s in srcs:
                er += grus.bfs(g, s, copy=False)[1].edges_reached
            res["cus_%d" % n] = round(er / (time.perf_counter() - t0) / 1e9, 1)
    finally:
        ctx.set_cu_mask(None)
    log("bfs GTEPS by CU count: %s" % res)
    return res


def run_leg_child(name, args):
    """One extra leg (serving, cu, capped) in a child process with its own
    context and graph: the leg's worker threads, streams, CU-masked queues
    and device budget stay out of this process, and a crash there costs the
    leg, not the line.  Its log lines and timed windows are passed on."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        regions = os.path.join(tmp, "regions.json")
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--leg-child", name,
               "--scale", str(args.scale), "--regions-out", regions]
        limit = args.leg_limit - 10 if args.leg_limit > 20 else None
        try:
            # its log lines go straight to this process's stderr
            r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=limit)
        except subprocess.TimeoutExpired:
            return {"error": "%s leg gave no result within %d s" % (name, limit)}
        if os.path.exists(regions):
            with open(regions) as f:
                REGIONS.windows.extend(json.load(f))
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    if r.returncode != 0 or not lines:
        return {"error": "%s leg exited with status %d" % (name, r.returncode)}
    return json.loads(lines[-1])


def run_serving(grus):
    """Batched serving on RMAT s22 (w in [1,255]) vs a one-query-at-a-time loop.

    Queries: 50/50 BFS/SSSP, sources uniform over non-isolated vertices
    (seed 5).  Saturation: 4096 queries offered at once.  Latency: open-loop
    Poisson arrivals (seed 4) at 30%, 60% and 90% of the measured saturation
    rate, submit -> result-ready on the server's clock, 3400 queries each."""
    ctx = grus.Context(0)
    g = grus.Graph.rmat(ctx, 22, weights=(1, 255), wseed=3)
    off, _, _ = g.csr()
    nz = np.nonzero(np.diff(off.astype(np.int64)))[0]
    rng = np.random.default_rng(5)
    # saturation: N queries offered at once (16 SSSP batches per engine of
    # the default 4, so the last batches' stragglers weigh little)
    N, NP = 4096, 3400
    kinds = rng.integers(0, 2, max(N, NP))
    srcs = rng.choice(nz, max(N, NP))
    # serial: best single-query kernels, one at a time, digest per query;
    # one untimed query of each kind first (SSSP builds its binned layout on
    # its first call)
    grus.bfs(g, int(srcs[0]), copy=False)
    grus.sssp(g, int(srcs[0]), copy=False)
    ns = 128
    lat = []
    t0 = time.perf_counter()
    for i in range(ns):
        q0 = time.perf_counter()
        (grus.bfs if kinds[i] == 0 else grus.sssp)(g, int(srcs[i]), copy=False)
        g.digest()
        lat.append(time.perf_counter() - q0)
    serial_qps = ns / (time.perf_counter() - t0)
    res = {"serial_qps": round(serial_qps, 1),
           "serial_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
           "serial_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2)}

    def saturate(**opts):
        with grus.Server(g, **opts) as srv:
            t0 = time.perf_counter()
            for i in range(N):
                srv.submit(int(kinds[i]), int(srcs[i]))
            out = [srv.poll(300000) for _ in range(N)]
            tb = time.perf_counter() - t0
            st = srv.stats()
        assert all(r.status == 0 for r in out)
        return N / tb, st

    # the library's default policy; and lane-by-lane admission into 8
    # engines per kind
    REGIONS.begin("serving_rmat_s22_saturation_query")
    qps, sat_stats = saturate()
    REGIONS.end(N)
    qps_cont, _ = saturate(max_inflight_batches=8, admit_min_free=1)
    res["batched_qps"] = round(qps, 1)
    res["speedup_vs_serial"] = round(qps / serial_qps, 2)
    res["policy"] = ("4 engines per kind x 64 lanes (BFS 64 queries, SSSP 128); an idle engine "
                     "admits a batch, a running batch admits arrivals into free slots during "
                     "its first 2 steps; each query reported when its slot finishes; an idle "
                     "engine serves its kind's queue one query at a time with the single-source "
                     "kernels while that empties the queue within one batch time")
    res["continuous_8_engines_qps"] = round(qps_cont, 1)
    # engines take device state as batches need it (admission by memory)
    res["saturation_engines_started"] = int(sat_stats["engines_started"])
    res["saturation_state_GB"] = round(sat_stats["state_bytes"] / 1e9, 2)
    # one full SSSP batch alone (128 queries in 64 lanes) and one of 64:
    # the reference points for latency
    for nq in (64, 128):
        ssrc = [int(x) for x in srcs[:nq]]
        grus.sssp_multi(g, ssrc, copy=False)
        res["sssp_batch%d_ms" % nq] = round(min(grus.sssp_multi(g, ssrc, copy=False)[1].device_ms
                                                for _ in range(3)), 2)
    bsrc = [int(x) for x in srcs[64:128]]
    grus.bfs_multi(g, bsrc, copy=False)
    res["bfs_batch64_ms"] = round(min(grus.bfs_multi(g, bsrc, copy=False)[1].device_ms
                                      for _ in range(3)), 2)
    sweep = []
    total = N * 2
    for frac in (0.3, 0.6, 0.9):
        rate = frac * qps
        gaps = np.random.default_rng(4).exponential(1.0 / rate, NP)
        with grus.Server(g) as srv:
            t0 = time.perf_counter()
            due = t0
            for i in range(NP):
                due += gaps[i]
                while time.perf_counter() < due:
                    pass
                srv.submit(int(kinds[i]), int(srcs[i]))
            out = [srv.poll(300000) for _ in range(NP)]
            st = srv.stats()
        assert all(r.status == 0 for r in out)
        lat = np.array([(r.done_us - r.submit_us) / 1e3 for r in out])
        sweep.append({"load": frac, "rate_qps": round(rate, 1),
                      "p50_ms": round(float(np.percentile(lat, 50)), 2),
                      "p99_ms": round(float(np.percentile(lat, 99)), 2),
                      "engines_started": int(st["engines_started"])})
        total += NP
    mid = sweep[1]
    res.update({"poisson_rate_qps": mid["rate_qps"], "p50_ms": mid["p50_ms"],
                "p99_ms": mid["p99_ms"], "latency_sweep": sweep, "queries": total,
                "mix": "50/50 BFS/SSSP, sources uniform over non-isolated vertices"})
    log("serving: serial %.1f qps, batched %.1f qps (x%.2f; lane-by-lane admission %.1f qps), "
        "p50/p99 ms at 30/60/90%% load: %s; one full batch: SSSP (128) %.1f ms, BFS (64) %.1f ms" %
        (serial_qps, qps, qps / serial_qps, qps_cont,
         " ".join("%.0f/%.0f" % (x["p50_ms"], x["p99_ms"]) for x in sweep),
         res["sssp_batch128_ms"], res["bfs_batch64_ms"]))
    g.close()
    ctx.close()
    return res


def run_capped(grus, ctx):
    """BFS + PageRank (2 iterations) on RMAT s27: resident in HBM vs the whole
    device capped at 8 GiB (graph part, app workspaces and every other buffer
    of the library; most edges then stream from pinned host memory).  The
    peak of the bytes the library held on the device during the capped apps
    is reported beside the cap."""
    cap = 8 << 30
    res = {"device_cap_bytes": cap}
    for mode in ("resident", "capped"):
        pl = grus.place("capped", device_budget_bytes=cap) if mode == "capped" else None
        g = grus.Graph.rmat(ctx, 27, placement=pl)
        try:
            held, _ = ctx.device_memory(reset_peak=True)
            src = g.info.max_degree_vertex
            grus.bfs(g, src, copy=False)
            REGIONS.begin("bfs_rmat_s27_%s" % mode)
            _, st = grus.bfs(g, src, copy=False)
            REGIONS.end(1)
            digest = g.digest()
            # top-down only: rows read where they live (zero-copy over the
            # link when capped), and the two-stage streaming BFS (rows
            # gathered into a device buffer per level, then expanded)
            REGIONS.begin("bfs_push_rmat_s27_%s" % mode)
            _, sp_push = grus.bfs(g, src, "push", copy=False)
            REGIONS.end(1)
            push_digest = g.digest()
            grus.bfs_2s(g, src, copy=False)
            REGIONS.begin("bfs_2s_rmat_s27_%s" % mode)
            _, s2 = grus.bfs_2s(g, src, copy=False)
            REGIONS.end(1)
            s2_digest = g.digest()
            REGIONS.begin("pagerank_2it_rmat_s27_%s" % mode)
            _, sp = grus.pagerank(g, 0.85, 2, copy=False)
            REGIONS.end(1)
            _, peak = ctx.device_memory()
            B = 2 * (8 * (g.V + 1) + 8 * g.E + 12 * g.V)
            res[mode] = {"bfs_ms": round(st.device_ms, 2),
                         "bfs_GTEPS": round(st.edges_reached / st.device_ms / 1e6, 2),
                         "bfs_push_only_ms": round(sp_push.device_ms, 2),
                         "bfs_2s_ms": round(s2.device_ms, 2),
                         "bfs_2s_stage_edges": int(s2.reserved[0]),
                         "push_and_2s_levels_equal": push_digest == digest and s2_digest == digest,
                         "pagerank_2it_ms": round(sp.device_ms, 1),
                         "pagerank_GBps": round(B / sp.device_ms / 1e6, 1),
                         "hbm_graph_GB": round(g.info.device_bytes / 1e9, 2),
                         "host_edges_GB": round(g.info.host_bytes / 1e9, 2),
                         "device_bytes_after_placement": int(held),
                         "device_peak_bytes_during_apps": int(peak),
                         "bfs_digest": digest[1]}
            if mode == "resident":
                res["V"], res["E"] = int(g.V), int(g.E)
        finally:
            g.close()
            ctx.set_device_budget(0)
    res["capped_within_cap"] = res["capped"]["device_peak_bytes_during_apps"] <= cap
    res["bfs_levels_identical"] = res["resident"].pop("bfs_digest") == res["capped"].pop("bfs_digest")
    log("capped s27: %s" % res)
    return res


if __name__ == "__main__":
    main()
