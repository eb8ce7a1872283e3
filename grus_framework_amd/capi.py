"""ctypes declarations of include/grus_c.h (the C-ABI of libgrus_hip.so).

The library is loaded from this package directory only; there is no fallback
implementation, so a missing build fails loudly.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrus_hip.so")

INFINIT = 0xFFFFFFFE

OK = 0
E_INVALID, E_NOMEM, E_HIP, E_IO, E_FORMAT = -1, -2, -3, -4, -5
E_BUDGET, E_TIMEOUT, E_UNSUPPORTED, E_NODEVICE, E_CLOSED = -6, -7, -8, -9, -10

PLACE_HBM, PLACE_MANAGED, PLACE_CAPPED = 0, 1, 2
DIR_AUTO, DIR_PUSH, DIR_PULL = 0, 1, 2
Q_BFS, Q_SSSP = 0, 1
TUNE_BFS, TUNE_SSSP, TUNE_PR = 0, 1, 2


class Place(C.Structure):
    _fields_ = [("mode", C.c_int), ("hbm_cap_bytes", C.c_uint64), ("prefetch", C.c_int),
                ("accessed_by", C.c_int), ("read_mostly", C.c_int),
                ("preferred_location", C.c_int), ("device_budget_bytes", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("iterations", C.c_uint32),
                ("pull_iterations", C.c_uint32), ("vertices_reached", C.c_uint64),
                ("edges_reached", C.c_uint64), ("kernel_launches", C.c_uint64),
                ("reserved", C.c_uint64 * 4)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class Tuning(C.Structure):
    _fields_ = [("bfs_alpha", C.c_double), ("bfs_beta", C.c_double),
                ("bu_thread_edges", C.c_uint32), ("sssp_delta", C.c_uint32),
                ("pr_layout", C.c_uint32), ("pr_bin_bits", C.c_uint32),
                ("sssp_pull_div", C.c_uint32), ("sssp_local_hops", C.c_uint32),
                ("pr_task_edges", C.c_uint32), ("sssp_pull_layout", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class GraphInfo(C.Structure):
    _fields_ = [("V", C.c_uint64), ("E", C.c_uint64), ("weighted", C.c_int),
                ("symmetric", C.c_int), ("placement", C.c_int),
                ("max_degree_vertex", C.c_uint32), ("max_degree", C.c_uint64),
                ("isolated", C.c_uint64), ("device_bytes", C.c_uint64),
                ("host_bytes", C.c_uint64)]


class CsrViewC(C.Structure):
    _fields_ = [("off", C.c_void_p), ("col", C.c_void_p), ("w", C.c_void_p),
                ("col_far", C.c_void_p), ("w_far", C.c_void_p), ("split", C.c_uint64),
                ("V", C.c_uint64), ("E", C.c_uint64)]


class ServerOpts(C.Structure):
    _fields_ = [("max_inflight_batches", C.c_int), ("max_batch", C.c_int),
                ("state_budget_bytes", C.c_uint64), ("copy_results", C.c_int),
                ("admit_min_free", C.c_int), ("admit_window", C.c_int), ("solo_max", C.c_int),
                ("idle_release_ms", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("engines_held", C.c_int * 2), ("state_bytes", C.c_uint64),
                ("engines_started", C.c_uint64), ("engines_released", C.c_uint64),
                ("memory_waits", C.c_uint64)]


class QueryResult(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("reached", C.c_uint64),
                ("checksum", C.c_uint64), ("submit_us", C.c_double), ("start_us", C.c_double),
                ("done_us", C.c_double)]


_vp = C.c_void_p
_u32, _u64, _i, _d, _f = C.c_uint32, C.c_uint64, C.c_int, C.c_double, C.c_float
_pp = C.POINTER(C.c_void_p)

# (name, restype, argtypes) for every function declared in include/grus_c.h
SIGNATURES = [
    ("grus_last_error", C.c_char_p, []),
    ("grus_version", C.c_char_p, []),
    ("grus_init", _i, [_i, _pp]),
    ("grus_fini", _i, [_vp]),
    ("grus_device_count", _i, [C.POINTER(_i)]),
    ("grus_get_tuning", _i, [_vp, C.POINTER(Tuning)]),
    ("grus_set_tuning", _i, [_vp, C.POINTER(Tuning)]),
    ("grus_set_cu_mask", _i, [_vp, C.POINTER(C.c_uint32), C.c_uint32]),
    ("grus_cu_count", _i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    ("grus_set_device_budget", _i, [_vp, _u64]),
    ("grus_device_memory", _i, [_vp, C.POINTER(_u64), C.POINTER(_u64), _i]),
    ("grus_gr_header", _i, [C.c_char_p, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_i)]),
    ("grus_gr_read", _i, [C.c_char_p, _vp, _vp, _vp]),
    ("grus_gr_write", _i, [C.c_char_p, _u64, _u64, _vp, _vp, _vp]),
    ("grus_graph_from_csr", _i, [_vp, _u64, _u64, _vp, _vp, _vp, _i, C.POINTER(Place), _pp]),
    ("grus_graph_load_gr", _i, [_vp, C.c_char_p, _i, _i, C.POINTER(Place), _pp]),
    ("grus_graph_gen_rmat", _i, [_vp, _i, _i, _d, _d, _d, _u64, _u32, _u32, _u64,
                                 C.POINTER(Place), _pp]),
    ("grus_graph_gen_grid", _i, [_vp, _u32, _u32, _u32, _u32, _u64, C.POINTER(Place), _pp]),
    ("grus_graph_info_get", _i, [_vp, C.POINTER(GraphInfo)]),
    ("grus_graph_copy_csr", _i, [_vp, _vp, _vp, _vp]),
    ("grus_graph_view", _i, [_vp, _i, C.POINTER(CsrViewC)]),
    ("grus_ctx_stream", _i, [_vp, _pp]),
    ("grus_graph_destroy", _i, [_vp]),
    ("grus_bfs", _i, [_vp, _u32, _i, _vp, C.POINTER(Stats)]),
    ("grus_bfs_2s", _i, [_vp, _u32, _u64, _vp, C.POINTER(Stats)]),
    ("grus_sssp", _i, [_vp, _u32, _vp, C.POINTER(Stats)]),
    ("grus_pagerank", _i, [_vp, _f, _i, _vp, C.POINTER(Stats)]),
    ("grus_cc", _i, [_vp, _i, _vp, C.POINTER(Stats)]),
    ("grus_cc_noalloc", _i, [_vp, _i, _vp, C.POINTER(Stats)]),
    ("grus_result_copy", _i, [_vp, _vp]),
    ("grus_result_digest", _i, [_vp, C.POINTER(_u64), C.POINTER(_u64)]),
    ("grus_bfs_multi", _i, [_vp, _vp, _i, _vp, C.POINTER(Stats)]),
    ("grus_sssp_multi", _i, [_vp, _vp, _i, _vp, C.POINTER(Stats)]),
    ("grus_autotune", _i, [_vp, _i, _vp, _i, C.POINTER(Tuning)]),
    ("grus_graph_signature", _i, [_vp, C.POINTER(_u64)]),
    ("grus_server_create", _i, [_vp, C.POINTER(ServerOpts), _pp]),
    ("grus_server_submit", _i, [_vp, _i, _u32, _vp, C.POINTER(_u64)]),
    ("grus_server_poll", _i, [_vp, C.POINTER(QueryResult), _i]),
    ("grus_server_destroy", _i, [_vp]),
    ("grus_server_stats_get", _i, [_vp, C.POINTER(ServerStats)]),
]

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libgrus_hip.so is not built (run __graft_entry__.build() or make -C "
                "grus_framework_amd/csrc); there is no fallback implementation")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class GrusError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc):
    if rc != OK:
        raise GrusError(rc, lib().grus_last_error().decode())
    return rc
