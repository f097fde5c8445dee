"""Vectorised numpy restatement of the algorithm of the reference's own CPU path, test.py:55-156.

Test infrastructure only.  It is an independent second formulation of the matcher:
  * vertex-major (dn x pn, de x pe) arrays where the C oracle is pattern-major;
  * prep of test.py:100-116 (first touch of a vertex takes max(fill, value), untouched vertices stay 0);
  * loop of test.py:122-156 (-inf initialised maxima, normaliser subtracted after the scatter, then max with the
    fill values; cnull = 0 as test.py:97, or with cnull=True the commented-out term of test.py:93-95).
It differs from src/main.cu only in the prep value of vertices without in-edges / out-edges (0 here, the fill value
there), so on graphs where every vertex has both, it must agree with the oracle.
"""
import numpy as np
from scipy.spatial.distance import cdist
from scipy.special import logsumexp


def normprob(x):
    """test.py:55-57: column log-softmax, max-shifted."""
    return x - logsumexp(x, axis=0, keepdims=True)


def run(data_v, data_e_feat, data_src, data_dst, patt_v, patt_e_feat, patt_src, patt_dst, n_iter=None, cnull=False):
    dv = data_v.shape[0]
    pe = patt_src.shape[0]
    n_iter = patt_v.shape[0] if n_iter is None else n_iter
    cv = cdist(data_v, patt_v)  # dn x pn
    mu = normprob(-cv)
    cv = normprob(cv)
    mu_max = mu.max(axis=0)
    v_bak_max = mu_max[patt_src]
    v_fwd_max = mu_max[patt_dst]
    ce = cdist(data_e_feat, patt_e_feat)  # de x pe
    xe = normprob(-ce)
    ce = normprob(ce)
    c_null = np.zeros(pe)
    if cnull:  # test.py:93-95
        c_null = np.sqrt((patt_e_feat ** 2).sum(axis=-1))
        c_null = np.maximum(c_null, ce.max(axis=0))
        c_null = normprob(c_null)

    # prep (test.py:100-116): touched vertices take max(fill, max over their edges); untouched stay 0.
    fwd_max = np.zeros((dv, pe))
    bak_max = np.zeros((dv, pe))
    fwd_seen = np.zeros(dv, bool)
    bak_seen = np.zeros(dv, bool)
    fwd_seen[data_dst] = True
    bak_seen[data_src] = True
    tmp = np.full((dv, pe), -np.inf)
    np.maximum.at(tmp, data_dst, xe)
    fwd_max[fwd_seen] = np.maximum(tmp[fwd_seen], v_bak_max[None, :])
    tmp = np.full((dv, pe), -np.inf)
    np.maximum.at(tmp, data_src, xe)
    bak_max[bak_seen] = np.maximum(tmp[bak_seen], v_fwd_max[None, :])

    for _ in range(n_iter):
        v_fwd = mu[:, patt_dst] - fwd_max
        v_bak = mu[:, patt_src] - bak_max
        v_fwd_max = v_fwd.max(axis=0)
        v_bak_max = v_bak.max(axis=0)
        e_bak = v_fwd[data_src] - ce
        e_fwd = v_bak[data_src] - ce
        e_bak_norm = logsumexp(e_bak, axis=0, keepdims=True)
        e_fwd_norm = logsumexp(e_fwd, axis=0, keepdims=True)
        fwd_max = np.full((dv, pe), -np.inf)
        bak_max = np.full((dv, pe), -np.inf)
        np.maximum.at(bak_max, data_src, e_bak)
        np.maximum.at(fwd_max, data_dst, e_fwd)
        fwd_max = np.maximum(fwd_max - e_fwd_norm, (v_bak_max - c_null)[None, :])
        bak_max = np.maximum(bak_max - e_bak_norm, (v_fwd_max - c_null)[None, :])
        mu = -cv
        for q in range(pe):
            mu[:, patt_dst[q]] += fwd_max[:, q]
            mu[:, patt_src[q]] += bak_max[:, q]
        mu = normprob(mu)
    return mu
