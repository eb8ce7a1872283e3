"""numpy front-end for liboracle.so (the serial CPU restatement + OpenMP baselines).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product (grus_framework_amd/) never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

INFINIT = 0xFFFFFFFE  # include/common.cuh:52

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_last_error.restype = C.c_char_p
        L.or_gr_header.argtypes = [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]
        L.or_gr_read.argtypes = [C.c_char_p, _u64p, _u32p, C.c_void_p]
        L.or_gen_rmat_count.restype = C.c_int64
        L.or_gen_rmat_count.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64]
        L.or_gen_rmat_fill.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, _u64p,
                                       _u32p, C.c_void_p]
        L.or_rmat_samples.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, C.c_uint64, _u32p, _u32p]
        L.or_edge_weight.restype = C.c_uint32
        L.or_edge_weight.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]
        L.or_gen_grid_count.restype = C.c_uint64
        L.or_gen_grid_count.argtypes = [C.c_uint32, C.c_uint32]
        L.or_gen_grid_fill.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_uint64, _u64p, _u32p, C.c_void_p]
        L.or_fnv64.restype = C.c_uint64
        L.or_fnv64.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.or_bfs.argtypes = [C.c_uint64, _u64p, _u32p, C.c_uint32, _u32p]
        L.or_bfs_certify.argtypes = [C.c_uint64, _u64p, _u32p, C.c_uint32, _u32p,
                                     C.POINTER(C.c_uint64)]
        L.or_sssp.argtypes = [C.c_uint64, _u64p, _u32p, C.c_void_p, C.c_uint32, _u32p]
        L.or_cc.argtypes = [C.c_uint64, _u64p, _u32p, _u32p]
        L.or_pagerank.argtypes = [C.c_uint64, _u64p, _u32p, _u64p, C.c_double, C.c_int, _f64p]
        L.or_transpose.argtypes = [C.c_uint64, _u64p, _u32p, C.c_void_p, _u64p, _u32p,
                                   C.c_void_p]
        L.bl_bfs.restype = C.c_double
        L.bl_bfs.argtypes = [C.c_uint64, _u64p, _u32p, C.c_uint32, _u32p, C.c_int]
        L.bl_sssp.restype = C.c_double
        L.bl_sssp.argtypes = [C.c_uint64, _u64p, _u32p, C.c_void_p, C.c_uint32, C.c_uint32,
                              _u32p, C.c_int]
        L.bl_pagerank.restype = C.c_double
        L.bl_pagerank.argtypes = [C.c_uint64, _u64p, _u32p, C.c_float, C.c_int, _f32p, C.c_int]
        L.bl_cc.restype = C.c_double
        L.bl_cc.argtypes = [C.c_uint64, _u64p, _u32p, _u32p, C.c_int]
        L.bl_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class CSR:
    """Host CSR: off u64[V+1], col u32[E], w u32[E] or None."""

    def __init__(self, off, col, w=None):
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.col = np.ascontiguousarray(col, dtype=np.uint32)
        self.w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        self.V = len(self.off) - 1
        self.E = int(self.off[-1])

    def degrees(self):
        return np.diff(self.off.astype(np.int64))

    def fnv(self):
        h = fnv64(self.off)
        h = fnv64(self.col, h)
        if self.w is not None:
            h = fnv64(self.w, h)
        return h


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().or_last_error().decode())


def fnv64(a, h=0):
    a = np.ascontiguousarray(a)
    return int(lib().or_fnv64(_ptr(a), a.nbytes, h))


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weights=None, wseed=0):
    L = lib()
    E = L.or_gen_rmat_count(scale, edge_factor, a, b, c, seed)
    if E < 0:
        raise RuntimeError(L.or_last_error().decode())
    off = np.zeros((1 << scale) + 1, np.uint64)
    col = np.zeros(max(E, 1), np.uint32)
    w = None
    wlo, whi = (weights if weights else (1, 1))
    if weights:
        w = np.zeros(max(E, 1), np.uint32)
    _check(L.or_gen_rmat_fill(scale, edge_factor, a, b, c, seed, wlo, whi, wseed, off, col,
                              _ptr(w)))
    return CSR(off, col[:E], None if w is None else w[:E])


def rmat_samples(scale, i0, i1, a=0.57, b=0.19, c=0.19, seed=1):
    u = np.zeros(i1 - i0, np.uint32)
    v = np.zeros(i1 - i0, np.uint32)
    lib().or_rmat_samples(scale, a, b, c, seed, i0, i1, u, v)
    return u, v


def edge_weight(u, v, wlo, whi, wseed):
    return int(lib().or_edge_weight(u, v, wlo, whi, wseed))


def gen_grid(rows, cols, weights=None, wseed=0):
    L = lib()
    E = int(L.or_gen_grid_count(rows, cols))
    off = np.zeros(rows * cols + 1, np.uint64)
    col = np.zeros(max(E, 1), np.uint32)
    w = np.zeros(max(E, 1), np.uint32) if weights else None
    wlo, whi = (weights if weights else (1, 1))
    L.or_gen_grid_fill(rows, cols, wlo, whi, wseed, off, col, _ptr(w))
    return CSR(off, col[:E], None if w is None else w[:E])


def read_gr(path, need_weight=True):
    L = lib()
    V, E, ety = C.c_uint64(), C.c_uint64(), C.c_uint64()
    _check(L.or_gr_header(path.encode(), C.byref(V), C.byref(E), C.byref(ety)))
    off = np.zeros(V.value + 1, np.uint64)
    col = np.zeros(max(E.value, 1), np.uint32)
    w = np.zeros(max(E.value, 1), np.uint32) if need_weight else None
    _check(L.or_gr_read(path.encode(), off, col, _ptr(w)))
    return CSR(off, col[:E.value], None if w is None else w[:E.value]), int(ety.value)


def bfs(g, src):
    out = np.zeros(g.V, np.uint32)
    lib().or_bfs(g.V, g.off, g.col, src, out)
    return out


def bfs_certify(g, src, level):
    """(code, vertex): 0 when `level` are the hop distances from src on the
    symmetric graph g (or_bfs_certify), else the failed condition."""
    bad = C.c_uint64()
    level = np.ascontiguousarray(level, np.uint32)
    code = lib().or_bfs_certify(g.V, g.off, g.col, src, level, C.byref(bad))
    return int(code), int(bad.value)


def sssp(g, src, unit=False):
    out = np.zeros(g.V, np.uint32)
    _check(lib().or_sssp(g.V, g.off, g.col, None if unit else _ptr(g.w), src, out))
    return out


def cc(g):
    out = np.zeros(g.V, np.uint32)
    lib().or_cc(g.V, g.off, g.col, out)
    return out


def transpose(g):
    t_off = np.zeros(g.V + 1, np.uint64)
    t_col = np.zeros(max(g.E, 1), np.uint32)
    t_w = np.zeros(max(g.E, 1), np.uint32) if g.w is not None else None
    lib().or_transpose(g.V, g.off, g.col, _ptr(g.w), t_off, t_col, _ptr(t_w))
    return CSR(t_off, t_col[:g.E], None if t_w is None else t_w[:g.E])


def pagerank(g, alpha=0.85, iters=20, csc=None):
    """FP64 ranks; csc = in-edge CSR (defaults to g: symmetric graphs)."""
    t = g if csc is None else csc
    out = np.zeros(g.V, np.float64)
    lib().or_pagerank(g.V, t.off, t.col, g.off, alpha, iters, out)
    return out


# ---- OpenMP baselines: return (result, milliseconds) ----
def bl_bfs(g, src, threads=0):
    out = np.zeros(g.V, np.uint32)
    ms = lib().bl_bfs(g.V, g.off, g.col, src, out, threads)
    return out, ms


def bl_sssp(g, src, delta=1000, threads=0):
    out = np.zeros(g.V, np.uint32)
    ms = lib().bl_sssp(g.V, g.off, g.col, _ptr(g.w), src, delta, out, threads)
    return out, ms


def bl_pagerank(g, alpha=0.85, iters=20, threads=0):
    out = np.zeros(g.V, np.float32)
    ms = lib().bl_pagerank(g.V, g.off, g.col, alpha, iters, out, threads)
    return out, ms


def bl_cc(g, threads=0):
    out = np.zeros(g.V, np.uint32)
    ms = lib().bl_cc(g.V, g.off, g.col, out, threads)
    return out, ms


def max_threads():
    return int(lib().bl_max_threads())
