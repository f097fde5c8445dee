"""Synthetic workloads: data graphs, features and planted patterns for the matcher's configurations.

All arrays are numpy, host-side: features float32 (row-major), ids int32.  Generation runs in the native library
(ac_gen_* in include/ac_mi355x.h), multi-threaded and deterministic in the seed.

Configurations (sizes as the reference's benchmark plan states them):
  small : directed Erdős–Rényi, 1,000 vertices / 20,000 edges, 13-dim features; pattern 10 vertices / 12 edges.
  s18   : R-MAT scale 18, edge factor 16, (a, b, c) = (.57, .19, .19), 32-dim; pattern 12 / 16.
  s22   : R-MAT scale 22, same parameters, 64-dim; pattern 16 / 24.
Patterns are connected subgraphs of the data graph, their features copied with N(0, 0.01^2) noise added.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib

STREAM_VFEAT = 10
STREAM_EFEAT = 11
STREAM_PV_NOISE = 20
STREAM_PE_NOISE = 21

RMAT_ABC = (0.57, 0.19, 0.19)


@dataclass
class Graph:
    vfeat: np.ndarray  # nV x dV float32
    efeat: np.ndarray  # nE x dE float32
    src: np.ndarray  # int32
    dst: np.ndarray  # int32

    @property
    def num_nodes(self) -> int:
        return int(self.vfeat.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.src.shape[0])


@dataclass
class Pattern(Graph):
    vmap: np.ndarray = field(default=None)  # pattern vertex -> planted data vertex
    emap: np.ndarray = field(default=None)  # pattern edge -> planted data edge


@dataclass
class Workload:
    name: str
    data: Graph
    patterns: list
    seed: int = 0

    @property
    def pattern(self) -> Pattern:
        return self.patterns[0]


def er_edges(n: int, m: int, seed: int):
    src = np.empty(m, np.int32)
    dst = np.empty(m, np.int32)
    _lib.check(_lib.lib().ac_gen_er(n, m, seed, src.ctypes.data, dst.ctypes.data))
    return src, dst


def rmat_edges(scale: int, edge_factor: int = 16, abc=RMAT_ABC, seed: int = 0):
    L = _lib.lib()
    ne = ctypes.c_int64(0)
    ps = ctypes.c_void_p()
    pd = ctypes.c_void_p()
    _lib.check(L.ac_gen_rmat(scale, edge_factor, abc[0], abc[1], abc[2], seed, ctypes.byref(ne), ctypes.byref(ps),
                             ctypes.byref(pd)))
    try:
        n = ne.value
        src = np.ctypeslib.as_array(ctypes.cast(ps, ctypes.POINTER(ctypes.c_int32)), shape=(max(n, 1),))[:n].copy()
        dst = np.ctypeslib.as_array(ctypes.cast(pd, ctypes.POINTER(ctypes.c_int32)), shape=(max(n, 1),))[:n].copy()
    finally:
        L.ac_free(ps)
        L.ac_free(pd)
    return src, dst


def normal(shape, seed: int, stream: int, sigma: float = 1.0) -> np.ndarray:
    out = np.empty(shape, np.float32)
    _lib.check(_lib.lib().ac_gen_normal(out.ctypes.data, out.size, seed, stream, sigma))
    return out


def planted_pattern(data: Graph, n_pv: int, n_pe: int, seed: int, sigma: float = 0.01) -> Pattern:
    ps = np.empty(n_pe, np.int32)
    pd = np.empty(n_pe, np.int32)
    vmap = np.empty(n_pv, np.int32)
    emap = np.empty(n_pe, np.int64)
    _lib.check(_lib.lib().ac_gen_pattern(data.num_nodes, data.num_edges, data.src.ctypes.data, data.dst.ctypes.data,
                                         n_pv, n_pe, seed, ps.ctypes.data, pd.ctypes.data, vmap.ctypes.data,
                                         emap.ctypes.data))
    vf = data.vfeat[vmap] + normal((n_pv, data.vfeat.shape[1]), seed, STREAM_PV_NOISE, sigma)
    ef = data.efeat[emap] + normal((n_pe, data.efeat.shape[1]), seed, STREAM_PE_NOISE, sigma)
    return Pattern(vf.astype(np.float32), ef.astype(np.float32), ps, pd, vmap=vmap, emap=emap)


def data_graph(src, dst, n: int, dv: int, de: int, seed: int) -> Graph:
    vf = normal((n, dv), seed, STREAM_VFEAT)
    ef = normal((len(src), de), seed, STREAM_EFEAT)
    return Graph(vf, ef, src, dst)


def small(seed: int = 0) -> Workload:
    """Erdős–Rényi 1,000 / 20,000, 13-dim; pattern 10 / 12."""
    src, dst = er_edges(1000, 20000, seed)
    g = data_graph(src, dst, 1000, 13, 13, seed)
    return Workload("small-er1000", g, [planted_pattern(g, 10, 12, seed)], seed)


def rmat(scale: int, dim: int, n_pv: int, n_pe: int, seed: int = 0, edge_factor: int = 16) -> Workload:
    src, dst = rmat_edges(scale, edge_factor, RMAT_ABC, seed)
    g = data_graph(src, dst, 1 << scale, dim, dim, seed)
    return Workload(f"rmat{scale}-d{dim}-p{n_pv}x{n_pe}", g, [planted_pattern(g, n_pv, n_pe, seed)], seed)


def s18(seed: int = 0) -> Workload:
    return rmat(18, 32, 12, 16, seed)


def s22(seed: int = 0) -> Workload:
    return rmat(22, 64, 16, 24, seed)


def planted_patterns(data: Graph, sizes, seeds, sigma: float = 0.01):
    """Planted patterns of the given (vertices, edges) sizes (ac_gen_patterns: one incidence build, planted in
    parallel); entries that could not be planted are None."""
    k = len(sizes)
    npv = np.array([s[0] for s in sizes], np.int32)
    npe = np.array([s[1] for s in sizes], np.int32)
    sd = np.array(seeds, np.uint64)
    voff = np.concatenate([[0], np.cumsum(npv)]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(npe)]).astype(np.int64)
    ps = np.empty(eoff[-1], np.int32)
    pd = np.empty(eoff[-1], np.int32)
    vmap = np.empty(voff[-1], np.int32)
    emap = np.empty(eoff[-1], np.int64)
    status = np.empty(k, np.int32)
    _lib.check(_lib.lib().ac_gen_patterns(data.num_nodes, data.num_edges, data.src.ctypes.data, data.dst.ctypes.data,
                                          k, npv.ctypes.data, npe.ctypes.data, sd.ctypes.data, ps.ctypes.data,
                                          pd.ctypes.data, vmap.ctypes.data, emap.ctypes.data, status.ctypes.data))
    out = []
    for i in range(k):
        if status[i]:
            out.append(None)
            continue
        vm = vmap[voff[i]:voff[i + 1]]
        em = emap[eoff[i]:eoff[i + 1]]
        nv, ne = int(npv[i]), int(npe[i])
        vf = data.vfeat[vm] + normal((nv, data.vfeat.shape[1]), int(sd[i]), STREAM_PV_NOISE, sigma)
        ef = data.efeat[em] + normal((ne, data.efeat.shape[1]), int(sd[i]), STREAM_PE_NOISE, sigma)
        out.append(Pattern(vf.astype(np.float32), ef.astype(np.float32), ps[eoff[i]:eoff[i + 1]].copy(),
                           pd[eoff[i]:eoff[i + 1]].copy(), vmap=vm.copy(), emap=em.copy()))
    return out


def uneven_patterns(data: Graph, count: int, seed: int, v_range=(4, 32)) -> list:
    """`count` planted patterns with V uniform in v_range and E uniform in [V, 2V] (the batched workload's shape).
    A pattern whose edge count cannot be planted is drawn again with a new seed and fewer edges, never fewer than V."""
    rng = np.random.default_rng(seed)
    sizes = []
    for _ in range(count):
        nv = int(rng.integers(v_range[0], v_range[1] + 1))
        sizes.append((nv, int(rng.integers(nv, 2 * nv + 1))))
    seeds = [seed * 100003 + i * 17 for i in range(count)]
    out = planted_patterns(data, sizes, seeds)
    for attempt in range(1, 8):
        miss = [i for i, p in enumerate(out) if p is None]
        if not miss:
            break
        sz = [(sizes[i][0], max(sizes[i][0], sizes[i][1] - attempt * max(1, sizes[i][0] // 4))) for i in miss]
        got = planted_patterns(data, sz, [seeds[i] + attempt for i in miss])
        for i, p in zip(miss, got):
            out[i] = p
    if any(p is None for p in out):
        raise _lib.ACError(_lib.AC_ERR_ARG, "uneven_patterns: some patterns could not be planted")
    return out


def batched(seed: int = 0, count: int = 256, scale: int = 20, dim: int = 32):
    """R-MAT scale 20, 32-dim features, with `count` uneven planted patterns: (data graph, patterns)."""
    src, dst = rmat_edges(scale, 16, RMAT_ABC, seed)
    g = data_graph(src, dst, 1 << scale, dim, dim, seed)
    return g, uneven_patterns(g, count, seed)


CONFIGS = {"small": small, "s18": s18, "s22": s22}
