"""ORACLE — test infrastructure only.

ctypes binding of ``oracle/ac_oracle.cpp``, the FP64 CPU restatement of the reference matcher
(src/main.cu:66-257 + src/ac.hxx:17-186).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, and only to check or to time the CPU baseline; the product package never
imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "liboracle.so")
SOURCE = os.path.join(HERE, "ac_oracle.cpp")

OR_COMPAT_ZERO_TAIL = 1
OR_LITERAL_LSM = 2
OR_CNULL = 4


def build(verbose: bool = False) -> str:
    """Compile the restatement with g++ (OpenMP) into oracle/_build/liboracle.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-Wall", SOURCE, "-o", tmp]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class _Graph(ctypes.Structure):
    _fields_ = [
        ("nV", ctypes.c_int64),
        ("dV", ctypes.c_int64),
        ("nE", ctypes.c_int64),
        ("dE", ctypes.c_int64),
        ("vfeat", ctypes.c_void_p),
        ("efeat", ctypes.c_void_p),
        ("src", ctypes.c_void_p),
        ("dst", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        for name in ("or_run_f64", "or_run_f32"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [P, P, ctypes.c_int, ctypes.c_uint, ctypes.c_int, P, P, P, P]
        L.or_run_w_f64.restype = ctypes.c_int
        L.or_run_w_f64.argtypes = [P, P, ctypes.c_int, ctypes.c_uint, ctypes.c_int, P, P, P]
        L.or_cdist.argtypes = [i64, i64, i64, P, P, P]
        L.or_cdist_w.argtypes = [i64, i64, i64, P, P, P, P]
        L.or_row_log_softmax.argtypes = [i64, i64, P, ctypes.c_int]
        L.or_row_max.argtypes = [i64, i64, P, P]
        L.or_edge_max_reduce.argtypes = [i64, i64, i64, P, P, P, P]
        L.or_compute_mu.argtypes = [i64, i64, i64, i64, P, P, P, P, P, P]
        L.or_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


@dataclass
class Graph64:
    """Graph in the oracle's types: f64 features, int64 ids (the reference's Real/Int, src/helpers.hxx:3-4)."""

    vfeat: np.ndarray  # nV x dV
    efeat: np.ndarray  # nE x dE
    src: np.ndarray
    dst: np.ndarray

    @staticmethod
    def of(vfeat, efeat, src, dst) -> "Graph64":
        v = np.ascontiguousarray(vfeat, dtype=np.float64)
        e = np.ascontiguousarray(efeat, dtype=np.float64)
        if e.ndim == 1:
            e = e.reshape(len(src), -1)
        return Graph64(v, e, np.ascontiguousarray(src, dtype=np.int64), np.ascontiguousarray(dst, dtype=np.int64))

    def c(self) -> _Graph:
        return _Graph(
            self.vfeat.shape[0],
            self.vfeat.shape[1],
            self.src.shape[0],
            self.efeat.shape[1] if self.efeat.ndim == 2 else 0,
            self.vfeat.ctypes.data,
            self.efeat.ctypes.data,
            self.src.ctypes.data,
            self.dst.ctypes.data,
        )


def _ptr(a):
    return None if a is None else a.ctypes.data


def run(data: Graph64, patt: Graph64, n_iter: int = -1, flags: int = 0, threads: int = 0, precision: str = "f64",
        want_state: bool = False):
    """Whole matcher.  Returns MU as a dn x pn float64 array (src/main.cu:247 order) and, with want_state, the
    final FMax/RMax (pe x dn) and the elapsed prep+loop+transpose seconds."""
    dn, pn, pe = data.vfeat.shape[0], patt.vfeat.shape[0], patt.src.shape[0]
    mu = np.empty((dn, pn), dtype=np.float64)
    fmax = np.empty((pe, dn), dtype=np.float64) if want_state else None
    rmax = np.empty((pe, dn), dtype=np.float64) if want_state else None
    el = (ctypes.c_double * 2)()
    D, Pg = data.c(), patt.c()
    fn = lib().or_run_f64 if precision == "f64" else lib().or_run_f32
    rc = fn(ctypes.byref(D), ctypes.byref(Pg), n_iter, flags, threads, _ptr(mu), _ptr(fmax), _ptr(rmax),
            el)
    if rc != 0:
        raise ValueError("oracle rejected the inputs (feature dims differ or empty graph)")
    if want_state:
        return mu, fmax, rmax, el[0] + el[1]
    return mu


def run_weighted(data: Graph64, patt: Graph64, wv=None, we=None, n_iter: int = -1, threads: int = 0):
    """Whole matcher (FP64) with per-dimension cdist weights wv (vertex features) and we (edge features);
    None = unweighted.  Returns MU, dn x pn."""
    dn, pn = data.vfeat.shape[0], patt.vfeat.shape[0]
    mu = np.empty((dn, pn), dtype=np.float64)
    wv = None if wv is None else np.ascontiguousarray(wv, np.float64)
    we = None if we is None else np.ascontiguousarray(we, np.float64)
    D, Pg = data.c(), patt.c()
    rc = lib().or_run_w_f64(ctypes.byref(D), ctypes.byref(Pg), n_iter, 0, threads, _ptr(wv), _ptr(we), mu.ctypes.data)
    if rc != 0:
        raise ValueError("oracle rejected the inputs (feature dims differ or empty graph)")
    return mu


def timed_run(data: Graph64, patt: Graph64, n_iter: int = -1, threads: int = 0, precision: str = "f64"):
    """Runs the restatement as the timed CPU baseline (Real = double for "f64", float for "f32"); returns
    (prep_seconds, loop_seconds, MU)."""
    dn, pn = data.vfeat.shape[0], patt.vfeat.shape[0]
    mu = np.empty((dn, pn), dtype=np.float64)
    el = (ctypes.c_double * 2)()
    D, Pg = data.c(), patt.c()
    fn = lib().or_run_f64 if precision == "f64" else lib().or_run_f32
    rc = fn(ctypes.byref(D), ctypes.byref(Pg), n_iter, 0, threads, mu.ctypes.data, None, None, el)
    if rc != 0:
        raise ValueError("oracle rejected the inputs")
    return el[0], el[1], mu


def cdist(a: np.ndarray, b: np.ndarray, w=None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.empty((a.shape[0], b.shape[0]), np.float64)
    if w is not None:
        w = np.ascontiguousarray(w, np.float64)
        lib().or_cdist_w(a.shape[0], b.shape[0], a.shape[1], a.ctypes.data, b.ctypes.data, w.ctypes.data,
                         out.ctypes.data)
        return out
    lib().or_cdist(a.shape[0], b.shape[0], a.shape[1], a.ctypes.data, b.ctypes.data, out.ctypes.data)
    return out


def row_log_softmax(x: np.ndarray, literal: bool = False) -> np.ndarray:
    x = np.array(x, dtype=np.float64, order="C", copy=True)
    lib().or_row_log_softmax(x.shape[0], x.shape[1], x.ctypes.data, int(literal))
    return x


def row_max(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(x.shape[0], np.float64)
    lib().or_row_max(x.shape[0], x.shape[1], x.ctypes.data, out.ctypes.data)
    return out


def edge_max_reduce(vymax, xe, nodes, n_col_out) -> np.ndarray:
    vymax = np.ascontiguousarray(vymax, np.float64)
    xe = np.ascontiguousarray(xe, np.float64)
    nodes = np.ascontiguousarray(nodes, np.int64)
    n_row, n_col_in = xe.shape
    out = np.empty((n_row, n_col_out), np.float64)
    lib().or_edge_max_reduce(n_col_in, n_col_out, n_row, vymax.ctypes.data, xe.ctypes.data, out.ctypes.data,
                             nodes.ctypes.data)
    return out


def compute_mu(cv, fmax, rmax, srcs, dsts) -> np.ndarray:
    cv = np.ascontiguousarray(cv, np.float64)
    fmax = np.ascontiguousarray(fmax, np.float64)
    rmax = np.ascontiguousarray(rmax, np.float64)
    srcs = np.ascontiguousarray(srcs, np.int64)
    dsts = np.ascontiguousarray(dsts, np.int64)
    n_row_out, n_col_out = cv.shape
    n_row_in, n_col_in = fmax.shape
    out = np.empty_like(cv)
    lib().or_compute_mu(n_col_in, n_row_in, n_col_out, n_row_out, cv.ctypes.data, fmax.ctypes.data,
                        rmax.ctypes.data, srcs.ctypes.data, dsts.ctypes.data, out.ctypes.data)
    return out


def max_threads() -> int:
    return lib().or_max_threads()
