"""ctypes binding of libac_mi355x.so (declarations in include/ac_mi355x.h).

The library is built in-tree by ``application_classification_amd.build``; loading fails loudly if it is missing,
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libac_mi355x.so")

AC_OK = 0
AC_ERR_ARG = 1
AC_ERR_OOM = 2
AC_ERR_HIP = 3
AC_ERR_STATE = 4
AC_ERR_IO = 5
AC_ERR_NODEV = 6
AC_DATA = 0
AC_PATTERN = 1
AC_OPT_WALK_BATCH = 1
AC_OPT_GRAPH = 2
AC_OPT_SLAB_EDGES = 3
AC_OPT_TAPE = 4
AC_OPT_PRUNE = 5
AC_OPT_CDIST_MFMA = 6
AC_OPT_PACK_COSTS = 7
AC_OPT_CNULL = 8
STAGES = ["prep", "vertex_a", "reduce_a", "edge_walk", "walk_fixup", "vertex_c", "reduce_c", "output"]

ERROR_NAMES = {
    AC_ERR_ARG: "AC_ERR_ARG",
    AC_ERR_OOM: "AC_ERR_OOM",
    AC_ERR_HIP: "AC_ERR_HIP",
    AC_ERR_STATE: "AC_ERR_STATE",
    AC_ERR_IO: "AC_ERR_IO",
    AC_ERR_NODEV: "AC_ERR_NODEV",
}


class ACError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {msg}")
        self.code = code


class HostGraph(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_int64),
        ("node_feat_dim", ctypes.c_int64),
        ("num_edges", ctypes.c_int64),
        ("edge_feat_dim", ctypes.c_int64),
        ("node_feats", ctypes.POINTER(ctypes.c_float)),
        ("edge_feats", ctypes.POINTER(ctypes.c_float)),
        ("srcs", ctypes.POINTER(ctypes.c_int32)),
        ("dsts", ctypes.POINTER(ctypes.c_int32)),
    ]


class Graph64(ctypes.Structure):
    """The reference's Graph at its own precision (uint64 sizes and ids, FP64 features)."""

    _fields_ = [
        ("num_nodes", ctypes.c_uint64),
        ("node_feat_dim", ctypes.c_uint64),
        ("num_edges", ctypes.c_uint64),
        ("edge_feat_dim", ctypes.c_uint64),
        ("node_feats", ctypes.POINTER(ctypes.c_double)),
        ("edge_feats", ctypes.POINTER(ctypes.c_double)),
        ("srcs", ctypes.POINTER(ctypes.c_uint64)),
        ("dsts", ctypes.POINTER(ctypes.c_uint64)),
    ]


P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
F = ctypes.c_float
D = ctypes.c_double

# name -> (restype, argtypes); every symbol declared in include/ac_mi355x.h.
SIGNATURES = {
    "ac_graph_load_bin": (I, [ctypes.c_char_p, ctypes.POINTER(HostGraph)]),
    "ac_graph_save_bin": (I, [ctypes.c_char_p, ctypes.POINTER(HostGraph)]),
    "ac_graph_load_csv": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(HostGraph)]),
    "ac_graph_save_csv": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(HostGraph)]),
    "ac_graph_free": (None, [ctypes.POINTER(HostGraph)]),
    "ac_graph64_load_bin": (I, [ctypes.c_char_p, ctypes.POINTER(Graph64)]),
    "ac_graph64_save_bin": (I, [ctypes.c_char_p, ctypes.POINTER(Graph64)]),
    "ac_graph64_load_csv": (I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Graph64)]),
    "ac_graph64_free": (None, [ctypes.POINTER(Graph64)]),
    "ac_set_graph64": (I, [P, I, ctypes.POINTER(Graph64)]),
    "ac_gen_er": (I, [I64, I64, U64, P, P]),
    "ac_gen_rmat": (I, [I, I, D, D, D, U64, ctypes.POINTER(I64), ctypes.POINTER(P), ctypes.POINTER(P)]),
    "ac_gen_normal": (I, [P, I64, U64, U64, F]),
    "ac_gen_pattern": (I, [I64, I64, P, P, I, I, U64, P, P, P, P]),
    "ac_gen_patterns": (I, [I64, I64, P, P, I, P, P, P, P, P, P, P, P]),
    "ac_free": (None, [P]),
    "ac_create": (I, [ctypes.POINTER(P), I]),
    "ac_destroy": (I, [P]),
    "ac_last_error": (ctypes.c_char_p, [P]),
    "ac_global_error": (ctypes.c_char_p, []),
    "ac_set_graph": (I, [P, I, I64, I64, P, I64, I64, P, P, P]),
    "ac_run": (I, [P, I]),
    "ac_run_async": (I, [P, I]),
    "ac_sync": (I, [P]),
    "ac_get_mu": (I, [P, P]),
    "ac_get_match": (I, [P, F, P]),
    "ac_get_topk": (I, [P, I, P, P]),
    "ac_get_sizes": (I, [P, P]),
    "ac_get_timing": (I, [P, ctypes.POINTER(D), ctypes.POINTER(D)]),
    "ac_get_cost_packing": (I, [P, P]),
    "ac_stream": (P, [P]),
    "ac_run_profiled": (I, [P, I, P]),
    "ac_set_option": (I, [P, I, I]),
    "ac_set_weights": (I, [P, P, P]),
    "ac_backward": (I, [P, P, P, P]),
    "ac_get_option": (I, [P, I, P]),
    "ac_batch_set_patterns": (I, [P, I, P, P, I64, P, I64, P, P, P]),
    "ac_batch_run": (I, [P, P]),
    "ac_batch_get_mu": (I, [P, I, P]),
    "ac_batch_info": (I, [P, P]),
    "ac_batch_timing": (I, [P, ctypes.POINTER(D), ctypes.POINTER(D)]),
    "ac_batch_autotune": (I, [P, I, P, P]),
    "ac_cdist": (I, [I64, I64, I64, P, P, P, P]),
    "ac_row_log_softmax": (I, [I64, I64, P, P]),
    "ac_row_log_softmax_static": (I, [I64, I64, P, P]),
    "ac_row_max": (I, [I64, I64, P, P, P]),
    "ac_edge_max_reduce": (I, [I64, I64, I64, P, P, P, P, P]),
    "ac_compute_mu": (I, [I64, I64, I64, I64, P, P, P, P, P, P, P]),
    "ac_transpose": (I, [P, P, I64, I64, P]),
}

_lib = None


def lib():
    """Load the library (once) and declare every exported signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (or python -m application_classification_amd.build)"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, ctx=None):
    if rc != AC_OK:
        if ctx is not None:
            msg = lib().ac_last_error(ctx)
        else:
            msg = lib().ac_global_error()
        raise ACError(rc, (msg or b"").decode(errors="replace"))
