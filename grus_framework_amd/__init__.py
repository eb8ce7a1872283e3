"""grus_framework_amd -- MI355X (gfx950) graph-traversal engine.

Python front-end over the C-ABI of libgrus_hip.so (include/grus_c.h).  The
engine itself is C++/HIP; this module only moves numpy arrays across the
boundary.  Names follow the reference's surface: graph loading
(graph_loader::Load, include/graph_loader.cuh:160), placement
(Set_Mem_Policy, include/graph.cuh:170), and the BFS / SSSP / PageRank / CC
apps (samples/*).
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import (DIR_AUTO, DIR_PULL, DIR_PUSH, INFINIT, PLACE_CAPPED, PLACE_HBM,
                   PLACE_MANAGED, Q_BFS, Q_SSSP, GrusError, Place, Stats, Tuning)

__all__ = ["Context", "Graph", "Server", "autotune", "bfs", "sssp", "pagerank", "cc", "bfs_multi",
           "sssp_multi", "gr_header", "gr_read", "gr_write", "device_count", "GrusError",
           "INFINIT", "Place", "Stats", "Tuning", "place"]

_DIRS = {"auto": DIR_AUTO, "push": DIR_PUSH, "pull": DIR_PULL}


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _lib():
    return capi.lib()


def device_count():
    n = C.c_int(0)
    capi.check(_lib().grus_device_count(C.byref(n)))
    return n.value


def place(mode="hbm", hbm_cap_bytes=0, pf=False, ab=False, rm=False, pl=False,
          device_budget_bytes=0):
    """Placement request; flag names follow include/app.cuh:27-31.
    device_budget_bytes (capped mode) caps the whole device footprint: graph,
    app workspaces and everything else the library holds on the device."""
    modes = {"hbm": PLACE_HBM, "managed": PLACE_MANAGED, "capped": PLACE_CAPPED}
    return Place(modes[mode], int(hbm_cap_bytes), int(pf), int(ab), int(rm), int(pl),
                 int(device_budget_bytes))


# ------------------------------------------------------------------ .gr I/O
def gr_header(path):
    V, E, wt = C.c_uint64(), C.c_uint64(), C.c_int()
    capi.check(_lib().grus_gr_header(path.encode(), C.byref(V), C.byref(E), C.byref(wt)))
    return V.value, E.value, bool(wt.value)


def gr_read(path, need_weight=True):
    V, E, _ = gr_header(path)
    off = np.zeros(V + 1, np.uint64)
    col = np.zeros(max(E, 1), np.uint32)
    w = np.zeros(max(E, 1), np.uint32) if need_weight else None
    capi.check(_lib().grus_gr_read(path.encode(), _p(off), _p(col), _p(w)))
    return off, col[:E], (None if w is None else w[:E])


def gr_write(path, off, col, w=None):
    off = np.ascontiguousarray(off, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    w = None if w is None else np.ascontiguousarray(w, np.uint32)
    capi.check(_lib().grus_gr_write(path.encode(), len(off) - 1, len(col), _p(off), _p(col),
                                     _p(w)))


# ------------------------------------------------------------------ context
class Context:
    """One device + one HIP stream (include/grus_c.h grus_init)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        capi.check(_lib().grus_init(device, C.byref(h)))
        self.h = h

    @property
    def tuning(self):
        t = Tuning()
        capi.check(_lib().grus_get_tuning(self.h, C.byref(t)))
        return t

    @tuning.setter
    def tuning(self, t):
        capi.check(_lib().grus_set_tuning(self.h, C.byref(t)))

    def set_device_budget(self, nbytes):
        """Cap the device bytes the library may hold (0 = no cap)."""
        capi.check(_lib().grus_set_device_budget(self.h, int(nbytes)))

    def device_memory(self, reset_peak=False):
        """(bytes the library holds on the device, peak since the last reset)."""
        held, peak = C.c_uint64(), C.c_uint64()
        capi.check(_lib().grus_device_memory(self.h, C.byref(held), C.byref(peak),
                                             int(reset_peak)))
        return held.value, peak.value

    @property
    def cu_count(self):
        """(CUs the context may use, CUs of the device)."""
        a, t = C.c_int(), C.c_int()
        capi.check(_lib().grus_cu_count(self.h, C.byref(a), C.byref(t)))
        return a.value, t.value

    def set_cu_mask(self, words=None):
        """Run on the CUs whose bits are set (32 per word); None: every CU.
        Co-location analogue of the reference's tools/sm_limit.cu."""
        if not words:
            capi.check(_lib().grus_set_cu_mask(self.h, None, 0))
            return
        arr = (C.c_uint32 * len(words))(*[int(w) & 0xFFFFFFFF for w in words])
        capi.check(_lib().grus_set_cu_mask(self.h, arr, len(words)))

    def set_cu_limit(self, n):
        """Run on CUs 0..n-1 of the mask numbering (n >= the device's CUs: all)."""
        total = self.cu_count[1]
        if n >= total:
            self.set_cu_mask(None)
            return
        if n < 1:
            raise ValueError("at least one CU")
        words = [0] * ((total + 31) // 32)
        for i in range(n):
            words[i // 32] |= 1 << (i % 32)
        self.set_cu_mask(words)

    def close(self):
        if self.h:
            capi.check(_lib().grus_fini(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Graph:
    """Device-resident CSR graph (graph_t<CSR>, include/graph.cuh:254)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle
        info = capi.GraphInfo()
        capi.check(_lib().grus_graph_info_get(self.h, C.byref(info)))
        self.info = info
        self.V = info.V
        self.E = info.E

    @classmethod
    def from_csr(cls, ctx, off, col, w=None, symmetric=False, placement=None):
        off = np.ascontiguousarray(off, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        w = None if w is None else np.ascontiguousarray(w, np.uint32)
        h = C.c_void_p()
        pl = placement if placement is not None else place()
        capi.check(_lib().grus_graph_from_csr(ctx.h, len(off) - 1, len(col), _p(off), _p(col),
                                              _p(w), int(symmetric), C.byref(pl), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def load_gr(cls, ctx, path, need_weight=True, symmetric=False, placement=None):
        h = C.c_void_p()
        pl = placement if placement is not None else place()
        capi.check(_lib().grus_graph_load_gr(ctx.h, path.encode(), int(need_weight),
                                             int(symmetric), C.byref(pl), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def rmat(cls, ctx, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weights=None,
             wseed=0, placement=None):
        wlo, whi = weights if weights else (0, 0)
        h = C.c_void_p()
        pl = placement if placement is not None else place()
        capi.check(_lib().grus_graph_gen_rmat(ctx.h, scale, edge_factor, a, b, c, seed, wlo, whi,
                                              wseed, C.byref(pl), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def grid(cls, ctx, rows, cols, weights=None, wseed=0, placement=None):
        wlo, whi = weights if weights else (0, 0)
        h = C.c_void_p()
        pl = placement if placement is not None else place()
        capi.check(_lib().grus_graph_gen_grid(ctx.h, rows, cols, wlo, whi, wseed, C.byref(pl),
                                              C.byref(h)))
        return cls(ctx, h)

    def csr(self):
        off = np.zeros(self.V + 1, np.uint64)
        col = np.zeros(max(self.E, 1), np.uint32)
        w = np.zeros(max(self.E, 1), np.uint32) if self.info.weighted else None
        capi.check(_lib().grus_graph_copy_csr(self.h, _p(off), _p(col), _p(w)))
        return off, col[:self.E], (None if w is None else w[:self.E])

    def result(self, dtype=np.uint32):
        out = np.zeros(self.V, dtype)
        capi.check(_lib().grus_result_copy(self.h, _p(out)))
        return out

    def digest(self):
        """(reached, checksum) of the last u32 result (grus_result_digest)."""
        r, c = C.c_uint64(), C.c_uint64()
        capi.check(_lib().grus_result_digest(self.h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def close(self):
        if self.h:
            capi.check(_lib().grus_graph_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ------------------------------------------------------------------ apps
def bfs(g, src, direction="auto", copy=True):
    """BFS levels (samples/bfs/bfs.cu); returns (levels or None, Stats)."""
    st = Stats()
    out = np.zeros(g.V, np.uint32) if copy else None
    capi.check(_lib().grus_bfs(g.h, src, _DIRS[direction], _p(out), C.byref(st)))
    return out, st


def bfs_2s(g, src, stage_edges=0, copy=True):
    """Two-stage top-down BFS (samples/bfs_2s/bfs_2s.cu): frontier rows are
    gathered into a device buffer of stage_edges edges, then expanded."""
    st = Stats()
    out = np.zeros(g.V, np.uint32) if copy else None
    capi.check(_lib().grus_bfs_2s(g.h, src, int(stage_edges), _p(out), C.byref(st)))
    return out, st


def sssp(g, src, copy=True):
    """SSSP distances (samples/sssp/sssp.cu); returns (dist or None, Stats)."""
    st = Stats()
    out = np.zeros(g.V, np.uint32) if copy else None
    capi.check(_lib().grus_sssp(g.h, src, _p(out), C.byref(st)))
    return out, st


def pagerank(g, alpha=0.85, iters=20, copy=True):
    """FP32 pull PageRank (samples/pagerank/pr.cu, fixed iterations)."""
    st = Stats()
    out = np.zeros(g.V, np.float32) if copy else None
    capi.check(_lib().grus_pagerank(g.h, alpha, iters, _p(out), C.byref(st)))
    return out, st


def cc(g, method="hook", copy=True):
    """Component labels = min vertex id (samples/cc/cc.cu)."""
    st = Stats()
    out = np.zeros(g.V, np.uint32) if copy else None
    m = {"hook": 0, "lp": 1}[method]
    capi.check(_lib().grus_cc(g.h, m, _p(out), C.byref(st)))
    return out, st


def bfs_multi(g, srcs, copy=True):
    srcs = np.ascontiguousarray(srcs, np.uint32)
    st = Stats()
    out = np.zeros((len(srcs), g.V), np.uint32) if copy else None
    capi.check(_lib().grus_bfs_multi(g.h, _p(srcs), len(srcs), _p(out), C.byref(st)))
    return out, st


def sssp_multi(g, srcs, copy=True):
    srcs = np.ascontiguousarray(srcs, np.uint32)
    st = Stats()
    out = np.zeros((len(srcs), g.V), np.uint32) if copy else None
    capi.check(_lib().grus_sssp_multi(g.h, _p(srcs), len(srcs), _p(out), C.byref(st)))
    return out, st


def autotune(g, kind="bfs", srcs=(), cache=None):
    """Pick the fastest thresholds for g (grus_autotune) and install them in
    g's context.  With `cache` (a JSON file path) the result is stored under
    the graph signature and reused on later calls."""
    import json
    import os
    kinds = {"bfs": capi.TUNE_BFS, "sssp": capi.TUNE_SSSP, "pr": capi.TUNE_PR}
    fields = {"bfs": ("bfs_alpha", "bfs_beta", "bu_thread_edges"),
              "sssp": ("sssp_delta", "sssp_pull_div", "sssp_local_hops"),
              "pr": ("pr_layout", "pr_bin_bits")}[kind]
    sig = C.c_uint64()
    capi.check(_lib().grus_graph_signature(g.h, C.byref(sig)))
    key = "%s:%016x" % (kind, sig.value)
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
    t = g.ctx.tuning
    if key in store:
        for name, val in store[key].items():
            setattr(t, name, val)
        g.ctx.tuning = t
        return t
    srcs = np.ascontiguousarray(srcs, np.uint32)
    capi.check(_lib().grus_autotune(g.h, kinds[kind], _p(srcs) if len(srcs) else None, len(srcs),
                                    C.byref(t)))
    if cache:
        store[key] = {n: getattr(t, n) for n in fields}  # only what this kind tunes
        with open(cache, "w") as f:
            json.dump(store, f, indent=1)
    return t


class Server:
    """Continuous-batching query server (grus_server_*)."""

    def __init__(self, g, max_inflight_batches=0, max_batch=0, state_budget_bytes=0,
                 copy_results=False, admit_min_free=0, admit_window=0, solo_max=0,
                 idle_release_ms=0):
        """Zero for max_inflight_batches, max_batch, admit_min_free or
        admit_window picks the library default (4 engines per kind, 64 lanes,
        a batch admits arrivals during its first 2 steps, later only when
        idle; waiting queries of a kind are served one at a time with the
        single-source kernels while that would empty the queue within one
        batch's time); solo_max > 0 fixes that queue length instead,
        admit_window < 0 turns the window off, solo_max < 0 the single-query
        path.  state_budget_bytes caps the device bytes of all engines' state:
        engines beyond the first of each kind take state only when it fits,
        and give it back after idle_release_ms (0 = 1000 under a cap or a
        device budget, else never; < 0 = never)."""
        self.g = g
        opts = capi.ServerOpts(max_inflight_batches, max_batch, state_budget_bytes,
                               int(copy_results), admit_min_free, admit_window, solo_max,
                               idle_release_ms)
        h = C.c_void_p()
        capi.check(_lib().grus_server_create(g.h, C.byref(opts), C.byref(h)))
        self.h = h
        self._outs = {}  # query id -> caller's result array, held until polled

    def submit(self, kind, src, out=None):
        """Queue one query.  `out`, when given, receives the V-entry result
        (copy_results servers): it must be a writable C-contiguous uint32
        array of at least V entries, and it is kept alive here until poll()
        returns its query, because a worker thread writes it asynchronously."""
        if out is not None:
            if not isinstance(out, np.ndarray) or out.dtype != np.uint32:
                raise TypeError("out must be a numpy uint32 array")
            if not out.flags.c_contiguous or not out.flags.writeable:
                raise ValueError("out must be writable and C-contiguous")
            if out.size < self.g.V:
                raise ValueError("out holds %d entries, the graph has %d vertices"
                                 % (out.size, self.g.V))
        qid = C.c_uint64()
        capi.check(_lib().grus_server_submit(self.h, kind, src, _p(out), C.byref(qid)))
        if out is not None:
            self._outs[qid.value] = out
        return qid.value

    def stats(self):
        """Memory-based admission counters (grus_server_stats)."""
        st = capi.ServerStats()
        capi.check(_lib().grus_server_stats_get(self.h, C.byref(st)))
        return {"engines_held": (st.engines_held[0], st.engines_held[1]),
                "state_bytes": st.state_bytes, "engines_started": st.engines_started,
                "engines_released": st.engines_released, "memory_waits": st.memory_waits}

    def poll(self, timeout_ms=-1):
        r = capi.QueryResult()
        capi.check(_lib().grus_server_poll(self.h, C.byref(r), timeout_ms))
        self._outs.pop(r.id, None)
        return r

    def close(self):
        if self.h:
            # the destroy joins the workers: no result is written afterwards
            capi.check(_lib().grus_server_destroy(self.h))
            self.h = None
            self._outs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
