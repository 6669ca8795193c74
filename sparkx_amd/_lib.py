"""ctypes binding of libsparkx_hb.so (the C ABI in include/sparkx_hb.h and
include/sparkx_hb_graphgen.h).  Loading fails loudly when the library is
missing: there is no CPU fallback for the HyperBall engine."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SPARKX_HB_LIB selects another build of the same ABI (e.g. an experimental one)
LIB_PATH = os.environ.get("SPARKX_HB_LIB") or os.path.join(HERE, "libsparkx_hb.so")

HB_OK, HB_EINVAL, HB_ENOMEM, HB_EHIP, HB_ESTATE, HB_ERANGE = 0, -1, -2, -3, -4, -5
HB_FLAG_DENSE, HB_FLAG_GRAPH, HB_FLAG_NF, HB_FLAG_TIMING, HB_FLAG_KTIMING = 0x1, 0x2, 0x4, 0x8, 0x10
HB_REGS_REFERENCE, HB_REGS_LATEST = 0, 1
HB_LAYOUT_INPUT, HB_LAYOUT_DEGREE, HB_LAYOUT_CHAIN, HB_LAYOUT_AUTO = 0, 1, 2, 3
HB_CLOSENESS, HB_HARMONIC, HB_LIN = 0, 1, 2
HB_STATS_MAX_ITERS = 4096
HB_NKERNELS = 5
KERNEL_SLOTS = ("union_bin0", "union_bin1", "union_bin2", "hub_chunk", "hub_merge")

# every extern "C" symbol the headers declare
EXPORTS = (
    "hb_create", "hb_destroy", "hb_last_error", "hb_build_digest", "hb_stage_csr", "hb_run",
    "hb_read_registers", "hb_read_estimates", "hb_read_accumulators", "hb_centrality",
    "hb_neighbourhood_function", "hb_set_schedule", "hb_num_nodes", "hb_num_arcs",
    "hb_graph_er", "hb_graph_rmat", "hb_graph_kmer", "hb_graph_nodes", "hb_graph_arcs",
    "hb_graph_offsets", "hb_graph_neighbours", "hb_graph_free",
    "hb_batch_create", "hb_batch_destroy", "hb_batch_set_schedule", "hb_batch_run",
    "hb_batch_graph_stats", "hb_batch_read_registers", "hb_batch_read_estimates",
    "hb_batch_read_accumulators", "hb_autotune", "hb_batch_autotune",
    "hb_set_snapshots", "hb_read_snapshots", "hb_mlp_create", "hb_mlp_destroy", "hb_mlp_set_params",
    "hb_mlp_get_params", "hb_mlp_gradient", "hb_mlp_train", "hb_mlp_predict", "hb_surrogate",
    "hb_walk_features", "hb_surrogate_walks",
    "hb_exact_bfs", "hb_exact_centrality", "hb_eval_ranks", "hb_eval_metrics", "hb_eval_metrics_noalloc", "hb_bias_windows", "hb_set_layout",
    "hb_stage_mtx", "hb_stage_coo", "hb_read_csr", "hb_run_many", "hb_result_info", "hb_restore",
)


class HbMtxInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("nnz", C.c_uint64), ("entries", C.c_uint64), ("arcs", C.c_uint64),
                ("err_line", C.c_uint64), ("symmetric", C.c_int32), ("err_code", C.c_int32),
                ("ms_device", C.c_double)]


HB_COO_SYMMETRIZE, HB_COO_DEDUP, HB_COO_NO_LOOPS = 0x1, 0x2, 0x4


class HbGraphDesc(C.Structure):
    _fields_ = [
        ("offsets", C.c_void_p),
        ("off_bytes", C.c_int),
        ("n", C.c_uint64),
        ("nbrs", C.c_void_p),
        ("nbr_bytes", C.c_int),
        ("arcs", C.c_uint64),
        ("p", C.c_int),
        ("max_iters", C.c_int),
    ]


class HbBatchStats(C.Structure):
    _fields_ = [
        ("launches", C.c_int32),
        ("iterations_max", C.c_int32),
        ("iterations_total", C.c_uint64),
        ("arcs_gathered", C.c_uint64),
        ("tiles", C.c_uint32),
        ("ms_total", C.c_double),
        ("ms_init", C.c_double),
        ("nodes_recomputed", C.c_uint64),
        ("alg_bytes", C.c_uint64),
    ]


class HbStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("final_parity_latest", C.c_int32),
        ("p", C.c_int32),
        ("ms_total", C.c_double),
        ("ms_init", C.c_double),
        ("arcs_gathered", C.c_uint64),
        ("nodes_recomputed", C.c_uint64),
        ("increases", C.c_uint64 * HB_STATS_MAX_ITERS),
        ("iter_arcs", C.c_uint64 * HB_STATS_MAX_ITERS),
        ("iter_nodes", C.c_uint64 * HB_STATS_MAX_ITERS),
        ("iter_ms", C.c_float * HB_STATS_MAX_ITERS),
        ("nf", C.c_double * HB_STATS_MAX_ITERS),
        ("kernel_launches", C.c_uint64 * HB_NKERNELS),
        ("kernel_ms", C.c_double * HB_NKERNELS),
        ("kernel_arcs", C.c_uint64 * HB_NKERNELS),
        ("kernel_nodes", C.c_uint64 * HB_NKERNELS),
        ("kernel_items", C.c_uint32 * HB_NKERNELS),
        ("kernel_team", C.c_uint32 * HB_NKERNELS),
        ("kernel_first_arcs", C.c_uint64 * HB_NKERNELS),
        ("kernel_first_nodes", C.c_uint64 * HB_NKERNELS),
        ("push_iterations", C.c_int32),
        ("symmetric", C.c_int32),
        ("kernel_gathered", C.c_uint64 * HB_NKERNELS),
        ("layout", C.c_int32),
        ("ms_layout", C.c_float),
    ]


class HyperBallError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m sparkx_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, u8p, f32p, f64p = C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_float), C.POINTER(C.c_double)
    u32p = C.POINTER(C.c_uint32)
    u64p = C.POINTER(C.c_uint64)
    sig = {
        "hb_create": (C.c_int, [C.POINTER(vp), C.c_int]),
        "hb_destroy": (None, [vp]),
        "hb_last_error": (C.c_char_p, []),
        "hb_build_digest": (C.c_char_p, []),
        "hb_stage_csr": (C.c_int, [vp, C.c_void_p, C.c_int, C.c_uint64, C.c_void_p, C.c_int, C.c_uint64]),
        "hb_run": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint32, C.POINTER(HbStats)]),
        "hb_read_registers": (C.c_int, [vp, C.c_int, u8p]),
        "hb_read_estimates": (C.c_int, [vp, C.c_int, f32p]),
        "hb_read_accumulators": (C.c_int, [vp, f64p, f64p]),
        "hb_centrality": (C.c_int, [vp, C.c_int, C.c_int, f64p]),
        "hb_neighbourhood_function": (C.c_int, [vp, f64p, C.c_int]),
        "hb_set_schedule": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
        "hb_num_nodes": (C.c_uint64, [vp]),
        "hb_num_arcs": (C.c_uint64, [vp]),
        "hb_graph_er": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]),
        "hb_graph_rmat": (C.c_int, [C.POINTER(vp), C.c_int, C.c_uint64, C.c_double, C.c_double,
                                    C.c_double, C.c_uint64, C.c_int]),
        "hb_graph_kmer": (C.c_int, [C.POINTER(vp), C.c_uint64, C.c_int, C.c_double, C.c_int,
                                    C.c_uint64, C.c_int]),
        "hb_graph_nodes": (C.c_uint64, [vp]),
        "hb_graph_arcs": (C.c_uint64, [vp]),
        "hb_graph_offsets": (u32p, [vp]),
        "hb_graph_neighbours": (u32p, [vp]),
        "hb_graph_free": (None, [vp]),
        "hb_batch_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(HbGraphDesc), C.c_int]),
        "hb_batch_destroy": (None, [vp]),
        "hb_batch_set_schedule": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32]),
        "hb_batch_run": (C.c_int, [vp, C.c_uint32, C.POINTER(HbBatchStats)]),
        "hb_batch_graph_stats": (C.c_int, [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_uint64), C.c_int]),
        "hb_batch_read_registers": (C.c_int, [vp, C.c_int, C.c_int, u8p]),
        "hb_batch_read_estimates": (C.c_int, [vp, C.c_int, C.c_int, f32p]),
        "hb_batch_read_accumulators": (C.c_int, [vp, C.c_int, f64p, f64p]),
        "hb_autotune": (C.c_int, [vp, C.c_int, C.c_int, u32p, u32p, u32p, f64p]),
        "hb_batch_autotune": (C.c_int, [vp, u32p, u32p, u32p, f64p]),
        "hb_set_snapshots": (C.c_int, [vp, C.c_int]),
        "hb_read_snapshots": (C.c_int, [vp, f32p]),
        "hb_mlp_create": (C.c_int, [C.POINTER(vp), C.c_int, C.c_uint64]),
        "hb_mlp_destroy": (None, [vp]),
        "hb_mlp_set_params": (C.c_int, [vp, f32p, f32p]),
        "hb_mlp_get_params": (C.c_int, [vp, f32p, f32p]),
        "hb_mlp_gradient": (C.c_int, [vp, f32p, f32p, C.c_uint64, f32p, C.POINTER(C.c_float)]),
        "hb_mlp_train": (C.c_int, [vp, f32p, f32p, C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_uint64, f32p]),
        "hb_mlp_predict": (C.c_int, [vp, f32p, C.c_uint64, f32p]),
        "hb_surrogate": (C.c_int, [vp, vp, C.c_int, f32p, f32p, C.POINTER(C.c_double)]),
        "hb_walk_features": (C.c_int, [vp, C.c_int, f32p, C.POINTER(C.c_double)]),
        "hb_surrogate_walks": (C.c_int, [vp, vp, C.c_int, f32p, f32p, C.POINTER(C.c_double)]),
        "hb_exact_bfs": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                   f64p]),
        "hb_exact_centrality": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint64, C.c_uint64, f64p]),
        "hb_eval_ranks": (C.c_int, [C.c_int, f64p, C.c_uint64, u32p]),
        "hb_eval_metrics": (C.c_int, [C.c_int, f64p, f64p, C.c_uint64, C.c_double, f64p]),
        "hb_eval_metrics_noalloc": (C.c_int, [C.c_int, f64p, f64p, C.c_uint64, C.c_double, f64p]),
        "hb_bias_windows": (C.c_int, [C.c_int, f64p, C.POINTER(C.c_float), C.POINTER(C.c_uint8)]),
        "hb_restore": (C.c_int, [vp, C.c_int, u8p, f64p, f64p]),
        "hb_set_layout": (C.c_int, [vp, C.c_int]),
        "hb_stage_mtx": (C.c_int, [vp, C.c_char_p, C.c_uint64, C.POINTER(HbMtxInfo)]),
        "hb_stage_coo": (C.c_int, [vp, C.c_uint64, u32p, u32p, C.c_uint64, C.c_uint32]),
        "hb_read_csr": (C.c_int, [vp, u64p, u64p]),
        "hb_run_many": (C.c_int, [C.POINTER(vp), C.POINTER(C.c_int), C.c_int, C.c_int, C.c_uint32,
                                  C.POINTER(HbStats), C.c_int]),
        "hb_result_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        msg = load().hb_last_error()
        raise HyperBallError(rc, msg.decode() if msg else "")
    return rc


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))
