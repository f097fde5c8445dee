"""Test workloads beyond the benchmark configurations (test infrastructure)."""
import os
from dataclasses import dataclass

import numpy as np

from application_classification_amd import workloads as W

REF_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_data")


@dataclass
class Case:
    name: str
    data: W.Graph
    pattern: W.Graph


def _read_rows(path):
    """The reference's CSV layout (prob2bin.py:11-12): header skipped, space separated."""
    return np.loadtxt(path, skiprows=1, dtype=np.float64)


def georgiy(seed: int = 0) -> Case:
    """The reference's georgiy sample: its 1,000 data vertices (13 features) and its 10-vertex / 21-edge pattern
    (17 edge features), with the missing data edges replaced by a directed Erdős–Rényi graph of 20,000 edges whose
    features follow the structure of the pattern's edge rows: 10 values in [0, 10), the constant 1.0, then the last
    three features of the source and destination vertices interleaved (verified on georgiyPattern.Edges.csv)."""
    dv = _read_rows(os.path.join(REF_DATA, "georgiyData.Vertex.csv"))
    pv = _read_rows(os.path.join(REF_DATA, "georgiyPattern.Vertex.csv"))
    pe = _read_rows(os.path.join(REF_DATA, "georgiyPattern.Edges.csv"))
    assert (dv[:, 0] == np.arange(len(dv))).all() and (pv[:, 0] == np.arange(len(pv))).all()
    dvf = dv[:, 1:].astype(np.float32)
    src, dst = W.er_edges(len(dv), 20000, seed)
    r = np.random.default_rng(seed)
    ef = np.empty((len(src), 17), np.float64)
    ef[:, :10] = r.uniform(0.0, 10.0, size=(len(src), 10))
    ef[:, 10] = 1.0
    ef[:, 11::2] = dvf[src][:, -3:]
    ef[:, 12::2] = dvf[dst][:, -3:]
    data = W.Graph(dvf, ef.astype(np.float32), src, dst)
    patt = W.Graph(pv[:, 1:].astype(np.float32), pe[:, 2:].astype(np.float32), pe[:, 0].astype(np.int32),
                   pe[:, 1].astype(np.int32))
    return Case("georgiy", data, patt)


def tiny_hand(seed: int = 11) -> Case:
    """Six data vertices with a duplicate edge, a self-loop, a vertex without in-edges (5 -> ...), a vertex without
    out-edges (3) and an isolated vertex (4 has neither); a three-vertex pattern with a self-loop edge and an
    odd number of edges (not a multiple of the float4 row width)."""
    r = np.random.default_rng(seed)
    src = np.array([0, 1, 2, 0, 2, 5, 1, 5, 0], np.int32)
    dst = np.array([1, 2, 0, 1, 2, 0, 3, 1, 3], np.int32)
    data = W.Graph(r.normal(size=(6, 4)).astype(np.float32), r.normal(size=(len(src), 5)).astype(np.float32), src,
                   dst)
    ps = np.array([0, 1, 2, 2, 0], np.int32)
    pd = np.array([1, 2, 0, 2, 2], np.int32)
    patt = W.Graph(r.normal(size=(3, 4)).astype(np.float32), r.normal(size=(len(ps), 5)).astype(np.float32), ps, pd)
    return Case("tiny_hand", data, patt)


def rmat_case(scale: int, dim: int, n_pv: int, n_pe: int, seed: int = 0) -> Case:
    w = W.rmat(scale, dim, n_pv, n_pe, seed)
    return Case(w.name, w.data, w.pattern)


def georgiy64(seed: int = 0) -> Case:
    """The georgiy case at the reference's own precision: the sample CSVs read as prob2bin.py reads them (FP64,
    pandas' decimal conversion, through ac_graph64_load_csv) and the data edges' features drawn in FP64 with the
    structure georgiy() gives them; nothing is rounded to FP32.  Arrays: features float64, ids int64."""
    from application_classification_amd import matcher
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        no_edges = os.path.join(tmp, "none.csv")
        open(no_edges, "w").write("0 2\n")
        dv = matcher.load_csv64(os.path.join(REF_DATA, "georgiyData.Vertex.csv"), no_edges)
    pg = matcher.load_csv64(os.path.join(REF_DATA, "georgiyPattern.Vertex.csv"),
                            os.path.join(REF_DATA, "georgiyPattern.Edges.csv"))
    src, dst = W.er_edges(dv.num_nodes, 20000, seed)
    r = np.random.default_rng(seed)
    ef = np.empty((len(src), 17), np.float64)
    ef[:, :10] = r.uniform(0.0, 10.0, size=(len(src), 10))
    ef[:, 10] = 1.0
    ef[:, 11::2] = dv.vfeat[src][:, -3:]
    ef[:, 12::2] = dv.vfeat[dst][:, -3:]
    data = W.Graph(dv.vfeat, ef, src.astype(np.int64), dst.astype(np.int64))
    patt = W.Graph(pg.vfeat, pg.efeat, pg.src.astype(np.int64), pg.dst.astype(np.int64))
    return Case("georgiy64", data, patt)
