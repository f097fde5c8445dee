"""Host-side mirror of the reference's driver and operator library, on top of the C ABI.

``Matcher`` is the driver (src/main.cu:66-257): load or hand over the two graphs, run prep + the loop, read MU.
``ac`` mirrors the ``ac::`` operator namespace (src/ac.hxx) over device buffers (torch.cuda tensors), with the
reference's names and argument meaning, for per-stage use and checks.

There is no CPU fallback: every call goes to libac_mi355x.so and raises ``ACError`` on failure.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (AC_DATA, AC_OPT_CNULL, AC_OPT_GRAPH, AC_OPT_PACK_COSTS, AC_OPT_PRUNE,  # noqa: F401
                   AC_OPT_SLAB_EDGES,
                   AC_OPT_TAPE, AC_OPT_WALK_BATCH, AC_PATTERN,
                   ACError, Graph64, HostGraph)


@dataclass
class GraphArrays:
    """A graph as host arrays: features float32 row-major, ids int32 (the reference's Graph, src/main.cu:8-19)."""

    vfeat: np.ndarray
    efeat: np.ndarray
    src: np.ndarray
    dst: np.ndarray

    @property
    def num_nodes(self) -> int:
        return int(self.vfeat.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.src.shape[0])


def _as_graph(g) -> GraphArrays:
    vf = np.ascontiguousarray(g.vfeat, dtype=np.float32)
    ef = np.ascontiguousarray(g.efeat, dtype=np.float32)
    if ef.ndim == 1:
        ef = ef.reshape(len(g.src), -1)
    return GraphArrays(vf, ef, np.ascontiguousarray(g.src, dtype=np.int32), np.ascontiguousarray(g.dst, dtype=np.int32))


def load_bin(path: str) -> GraphArrays:
    """Read a graph in the binary layout of prob2bin.py:46-64 (what src/main.cu:21-63 reads)."""
    L = _lib.lib()
    hg = HostGraph()
    _lib.check(L.ac_graph_load_bin(path.encode(), ctypes.byref(hg)))
    try:
        return _copy_host_graph(hg)
    finally:
        L.ac_graph_free(ctypes.byref(hg))


def load_csv(vertex_csv: str, edge_csv: str) -> GraphArrays:
    """Read the CSV pair that prob2bin.py:11-39 converts."""
    L = _lib.lib()
    hg = HostGraph()
    _lib.check(L.ac_graph_load_csv(vertex_csv.encode(), edge_csv.encode(), ctypes.byref(hg)))
    try:
        return _copy_host_graph(hg)
    finally:
        L.ac_graph_free(ctypes.byref(hg))


def _copy_host_graph(hg: HostGraph) -> GraphArrays:
    nV, dV, nE, dE = hg.num_nodes, hg.node_feat_dim, hg.num_edges, hg.edge_feat_dim

    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    return GraphArrays(
        arr(hg.node_feats, nV * dV, np.float32).reshape(nV, dV),
        arr(hg.edge_feats, nE * dE, np.float32).reshape(nE, dE),
        arr(hg.srcs, nE, np.int32),
        arr(hg.dsts, nE, np.int32),
    )


def _host_graph(g: GraphArrays) -> HostGraph:
    return HostGraph(
        g.num_nodes,
        g.vfeat.shape[1],
        g.num_edges,
        g.efeat.shape[1] if g.efeat.ndim == 2 else 0,
        g.vfeat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        g.efeat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        g.src.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        g.dst.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
    )


def save_bin(path: str, g) -> None:
    g = _as_graph(g)
    _lib.check(_lib.lib().ac_graph_save_bin(path.encode(), ctypes.byref(_host_graph(g))))


def save_csv(vertex_csv: str, edge_csv: str, g) -> None:
    g = _as_graph(g)
    _lib.check(_lib.lib().ac_graph_save_csv(vertex_csv.encode(), edge_csv.encode(), ctypes.byref(_host_graph(g))))


@dataclass
class Graph64Arrays:
    """The reference's Graph at its own precision (src/main.cu:8-19): features float64, ids uint64."""

    vfeat: np.ndarray
    efeat: np.ndarray
    src: np.ndarray
    dst: np.ndarray

    @property
    def num_nodes(self) -> int:
        return int(self.vfeat.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.src.shape[0])


def _copy_graph64(g: Graph64) -> Graph64Arrays:
    nV, dV, nE, dE = g.num_nodes, g.node_feat_dim, g.num_edges, g.edge_feat_dim

    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

    return Graph64Arrays(arr(g.node_feats, nV * dV, np.float64).reshape(nV, dV),
                         arr(g.edge_feats, nE * dE, np.float64).reshape(nE, dE),
                         arr(g.srcs, nE, np.uint64), arr(g.dsts, nE, np.uint64))


def _graph64(g) -> tuple:
    """(Graph64 struct, the arrays it points into) for any graph-like object, at FP64 / uint64."""
    vf = np.ascontiguousarray(g.vfeat, dtype=np.float64)
    ef = np.ascontiguousarray(g.efeat, dtype=np.float64)
    if ef.ndim == 1:
        ef = ef.reshape(len(g.src), -1)
    s = np.ascontiguousarray(g.src, dtype=np.uint64)
    d = np.ascontiguousarray(g.dst, dtype=np.uint64)
    st = Graph64(vf.shape[0], vf.shape[1], s.shape[0], ef.shape[1] if ef.ndim == 2 else 0,
                 vf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                 s.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), d.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return st, (vf, ef, s, d)


def load_bin64(path: str) -> Graph64Arrays:
    """Read a .bin at full precision (loadGraph, src/main.cu:21-47)."""
    L = _lib.lib()
    g = Graph64()
    _lib.check(L.ac_graph64_load_bin(path.encode(), ctypes.byref(g)))
    try:
        return _copy_graph64(g)
    finally:
        L.ac_graph64_free(ctypes.byref(g))


def load_csv64(vertex_csv: str, edge_csv: str) -> Graph64Arrays:
    """Read the CSV pair as prob2bin.py:11-39 does (pandas' decimal conversion), at full precision."""
    L = _lib.lib()
    g = Graph64()
    _lib.check(L.ac_graph64_load_csv(vertex_csv.encode(), edge_csv.encode(), ctypes.byref(g)))
    try:
        return _copy_graph64(g)
    finally:
        L.ac_graph64_free(ctypes.byref(g))


def save_bin64(path: str, g) -> None:
    """Write the .bin of prob2bin.py:46-64 from FP64 features."""
    st, keep = _graph64(g)
    _lib.check(_lib.lib().ac_graph64_save_bin(path.encode(), ctypes.byref(st)))
    del keep


class Matcher:
    """One matcher context on one GPU (ac_create / ac_set_graph / ac_run / ac_get_mu)."""

    def __init__(self, device: int = 0):
        self._L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(self._L.ac_create(ctypes.byref(h), device))
        self._h = h
        self._dn = self._pn = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.ac_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        _lib.check(rc, self._h)

    def _set(self, which: int, g) -> GraphArrays:
        g = _as_graph(g)
        dE = g.efeat.shape[1] if g.efeat.ndim == 2 else 0
        self._check(self._L.ac_set_graph(self._h, which, g.num_nodes, g.vfeat.shape[1], g.vfeat.ctypes.data,
                                         g.num_edges, dE, g.efeat.ctypes.data, g.src.ctypes.data, g.dst.ctypes.data))
        return g

    def set_data(self, g) -> None:
        g = self._set(AC_DATA, g)
        self._dn = g.num_nodes
        self._feat_dims = (g.vfeat.shape[1], g.efeat.shape[1] if g.efeat.ndim == 2 else 0)

    def set_graph64(self, which: int, g) -> None:
        """Hand over a graph at the reference's own precision (ac_set_graph64: FP64 features, uint64 ids)."""
        st, keep = _graph64(g)
        self._check(self._L.ac_set_graph64(self._h, which, ctypes.byref(st)))
        del keep
        if which == AC_DATA:
            self._dn = int(st.num_nodes)
            self._feat_dims = (int(st.node_feat_dim), int(st.edge_feat_dim))
        else:
            self._pn = int(st.num_nodes)

    def set_pattern(self, g) -> None:
        self._pn = self._set(AC_PATTERN, g).num_nodes

    def run(self, n_iter: int = -1) -> np.ndarray:
        """Prep + n_iter iterations (default: number of pattern vertices, src/main.cu:160); returns MU, dn x pn."""
        self._check(self._L.ac_run(self._h, n_iter))
        return self.mu()

    def run_async(self, n_iter: int = -1) -> None:
        self._check(self._L.ac_run_async(self._h, n_iter))

    def sync(self) -> None:
        self._check(self._L.ac_sync(self._h))

    def mu(self) -> np.ndarray:
        # Sized from the library's own record of the graphs: ac_get_mu writes dn x pn floats.
        dn, _, pn, _ = self.sizes()
        out = np.empty((dn, pn), np.float32)
        self._check(self._L.ac_get_mu(self._h, out.ctypes.data))
        return out

    def match(self, log_tau: float = float(np.log(0.5))) -> np.ndarray:
        dn, _, pn, _ = self.sizes()
        out = np.empty((dn, pn), np.uint8)
        self._check(self._L.ac_get_match(self._h, log_tau, out.ctypes.data))
        return out

    def topk(self, k: int):
        """(idx, val), each pn x k: per pattern vertex the k most likely data vertices (ac_get_topk)."""
        n, _, pn, _ = self.sizes()
        idx = np.empty((pn, k), np.int32)
        val = np.empty((pn, k), np.float32)
        self._check(self._L.ac_get_topk(self._h, k, idx.ctypes.data, val.ctypes.data))
        return idx, val

    def timing(self):
        """(prep_ms, loop_ms) of the last run, from HIP events on the context's stream."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self._L.ac_get_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def run_profiled(self, n_iter: int = -1) -> dict:
        """Runs like run() with a HIP event after each launch; returns device ms per stage (summed over iterations)."""
        out = (ctypes.c_double * len(_lib.STAGES))()
        self._check(self._L.ac_run_profiled(self._h, n_iter, out))
        return dict(zip(_lib.STAGES, list(out)))

    def sizes(self):
        out = (ctypes.c_int64 * 4)()
        self._check(self._L.ac_get_sizes(self._h, out))
        return tuple(out)

    def cost_packing(self):
        """(edge bundles, bundles whose costs the last run's walk read as 24-bit codes; AC_OPT_PACK_COSTS)."""
        out = (ctypes.c_int64 * 2)()
        self._check(self._L.ac_get_cost_packing(self._h, out))
        return tuple(out)

    @property
    def stream(self) -> int:
        return self._L.ac_stream(self._h) or 0

    # -- adjoint (ac_set_weights / ac_backward) -----------------------------------------------------------------
    def set_weights(self, wv=None, we=None) -> None:
        """Per-dimension cdist weights (None: unweighted); set after the data graph."""
        a = None if wv is None else np.ascontiguousarray(wv, np.float32)
        b = None if we is None else np.ascontiguousarray(we, np.float32)
        self._check(self._L.ac_set_weights(self._h, None if a is None else a.ctypes.data,
                                           None if b is None else b.ctypes.data))
        self._dims = (0 if a is None else a.size, 0 if b is None else b.size)

    def backward(self, grad_mu):
        """dL/dMU (dn x pn) of the last taped run (AC_OPT_TAPE > 0) -> (dL/dwv, dL/dwe)."""
        g = np.ascontiguousarray(grad_mu, np.float32)
        dn, _, pn, _ = self.sizes()
        if g.shape != (dn, pn):
            raise ValueError(f"grad_mu must be {dn} x {pn}")
        dv, de = self._feat_dims
        gwv = np.zeros(dv, np.float32)
        gwe = np.zeros(de, np.float32)
        self._check(self._L.ac_backward(self._h, g.ctypes.data, gwv.ctypes.data, gwe.ctypes.data))
        return gwv, gwe

    # -- options (ac_set_option) --------------------------------------------------------------------------------
    def set_option(self, option: int, value: int) -> None:
        self._check(self._L.ac_set_option(self._h, option, int(value)))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        self._check(self._L.ac_get_option(self._h, option, ctypes.byref(v)))
        return v.value

    # -- batched scoring (ac_batch_*) ----------------------------------------------------------------------------
    def batch_set_patterns(self, patterns) -> None:
        """Patterns (GraphArrays-like, local vertex ids) scored together against the data graph by batch_run."""
        pats = [_as_graph(p) for p in patterns]
        v_off = np.cumsum([0] + [p.num_nodes for p in pats]).astype(np.int32)
        e_off = np.cumsum([0] + [p.num_edges for p in pats]).astype(np.int32)
        dV = pats[0].vfeat.shape[1]
        dE = next((p.efeat.shape[1] for p in pats if p.efeat.ndim == 2 and p.num_edges), 0)
        vfeat = np.ascontiguousarray(np.concatenate([p.vfeat.reshape(-1, dV) for p in pats]), np.float32)
        efeat = np.ascontiguousarray(np.concatenate([p.efeat.reshape(-1, dE) for p in pats]), np.float32)
        src = np.ascontiguousarray(np.concatenate([p.src for p in pats]), np.int32)
        dst = np.ascontiguousarray(np.concatenate([p.dst for p in pats]), np.int32)
        self._check(self._L.ac_batch_set_patterns(self._h, len(pats), v_off.ctypes.data, e_off.ctypes.data, dV,
                                                  vfeat.ctypes.data, dE, efeat.ctypes.data, src.ctypes.data,
                                                  dst.ctypes.data))
        self._bpn = [p.num_nodes for p in pats]

    def batch_run(self, n_iter=None) -> None:
        """Runs every pattern (n_iter: per-pattern iteration counts, default each pattern's vertex count)."""
        arr = None if n_iter is None else np.ascontiguousarray(n_iter, np.int32)
        self._check(self._L.ac_batch_run(self._h, None if arr is None else arr.ctypes.data))

    def batch_mu(self, i: int) -> np.ndarray:
        out = np.empty((self.sizes()[0], self._bpn[i]), np.float32)
        self._check(self._L.ac_batch_get_mu(self._h, i, out.ctypes.data))
        return out

    def batch_info(self) -> dict:
        out = (ctypes.c_int64 * 4)()
        self._check(self._L.ac_batch_info(self._h, out))
        return {"patterns": out[0], "slabs": out[1], "updates": out[2], "updates_computed": out[3]}

    def batch_timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self._L.ac_batch_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def batch_autotune(self, probe_iters: int = 2):
        """Picks the slab size and walk batch for the current patterns; returns ((slab_edges, walk_batch), est_ms)."""
        ch = (ctypes.c_int32 * 2)()
        est = ctypes.c_double()
        self._check(self._L.ac_batch_autotune(self._h, probe_iters, ch, ctypes.byref(est)))
        return (ch[0], ch[1]), est.value


class ac:
    """The reference's operator library (src/ac.hxx) on device tensors (float32 values, int32 ids).

    Each method takes torch.cuda tensors (or raw device pointers as ints) with the reference's argument order and
    runs on ``stream`` (a hipStream_t as int; 0 = default stream).
    """

    @staticmethod
    def _p(x):
        return x if isinstance(x, int) else x.data_ptr()

    @staticmethod
    def cdist(n_a, n_b, dim, feats_a, feats_b, out, stream=0):
        """src/ac.hxx:33-56: out[j * n_b + i] = ||a_j - b_i||."""
        _lib.check(_lib.lib().ac_cdist(n_a, n_b, dim, ac._p(feats_a), ac._p(feats_b), ac._p(out), stream))

    @staticmethod
    def RowSoftmax2(n_row, n_col, x, stream=0):
        """src/ac.hxx:108-113: in-place row log-softmax."""
        _lib.check(_lib.lib().ac_row_log_softmax(n_row, n_col, ac._p(x), stream))

    RowSoftmax2_prealloc = RowSoftmax2

    @staticmethod
    def RowMax2(n_row, n_col, d_in, d_out, stream=0):
        """src/ac.hxx:58-75: per-row maximum."""
        _lib.check(_lib.lib().ac_row_max(n_row, n_col, ac._p(d_in), ac._p(d_out), stream))

    @staticmethod
    def EdgeMaxReduce2_t(n_col_in, n_col_out, n_row, VYMax, XE_t, XMax_t, nodes, stream=0):
        """src/ac.hxx:115-149: fill with VYMax then scatter-max XE_t by nodes."""
        _lib.check(_lib.lib().ac_edge_max_reduce(n_col_in, n_col_out, n_row, ac._p(VYMax), ac._p(XE_t),
                                                 ac._p(XMax_t), ac._p(nodes), stream))

    @staticmethod
    def ComputeMU2_t(n_col_in, n_row_in, n_col_out, n_row_out, CV_t, FMax_t, RMax_t, srcs, dsts, MU_t, stream=0):
        """src/ac.hxx:151-186: MU = -CV + scatter-add of FMax by dsts and RMax by srcs (ascending q)."""
        _lib.check(_lib.lib().ac_compute_mu(n_col_in, n_row_in, n_col_out, n_row_out, ac._p(CV_t), ac._p(FMax_t),
                                            ac._p(RMax_t), ac._p(srcs), ac._p(dsts), ac._p(MU_t), stream))

    @staticmethod
    def transpose(d_in, d_out, num_rows, num_cols, stream=0):
        """src/ac.hxx:17-31."""
        _lib.check(_lib.lib().ac_transpose(ac._p(d_in), ac._p(d_out), num_rows, num_cols, stream))


def run_matcher(data, pattern, n_iter: int = -1, device: int = 0) -> np.ndarray:
    """One-shot helper: MU (dn x pn) for a data graph and a pattern graph."""
    with Matcher(device) as m:
        m.set_data(data)
        m.set_pattern(pattern)
        return m.run(n_iter)
