"""The command-line driver (ac_main, the reference's `main data.bin pattern.bin > result`, src/main.cu:66-257): its
stdout is one "%e" line per MU value in dn x pn order, byte for byte the formatting of the library's own result, and
stderr carries the `elapsed=` line (src/main.cu:250)."""
import os
import subprocess

import numpy as np
import pytest

from application_classification_amd import workloads as W
from application_classification_amd.matcher import (AC_DATA, AC_PATTERN, Matcher, load_csv, save_bin, save_bin64,
                                                    save_csv)
from oracle import oracle as O
from tests import cases
from tests.conftest import allclose_report

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "application_classification_amd", "ac_main")


def expected_text(mu):
    return "".join("%e\n" % v for v in mu.astype(np.float64).ravel())


def run_cli(args):
    print(f"ac_main {' '.join(os.path.basename(a) for a in args)}", flush=True)
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=180)


def test_cli_bin_output_is_the_library_result(tmp_path):
    """R-MAT 17 with a 10-vertex pattern: 1.31 M values, more than one 2^20-value formatting chunk."""
    w = W.rmat(17, 8, 10, 12, seed=3)
    d, p = w.data, w.pattern
    db, pb = str(tmp_path / "data.bin"), str(tmp_path / "pattern.bin")
    save_bin(db, d)
    save_bin(pb, p)
    r = run_cli([db, pb])
    assert r.returncode == 0, r.stderr
    assert any(line.startswith("elapsed=") for line in r.stderr.splitlines())
    with Matcher(0) as m:
        m.set_data(d)
        m.set_pattern(p)
        mu = m.run()
    assert d.num_nodes * p.num_nodes > (1 << 20)
    assert r.stdout == expected_text(mu)


def test_cli_csv_quartet(tmp_path):
    """The CSV quartet and --iters; the library runs on the same files, read back, so both sides parse them alike."""
    w = W.small(0)
    paths = [str(tmp_path / n) for n in ("dv.csv", "de.csv", "pv.csv", "pe.csv")]
    save_csv(paths[0], paths[1], w.data)
    save_csv(paths[2], paths[3], w.pattern)
    r = run_cli(["--csv", *paths, "--iters", "3"])
    assert r.returncode == 0, r.stderr
    d, p = load_csv(paths[0], paths[1]), load_csv(paths[2], paths[3])
    with Matcher(0) as m:
        m.set_data(d)
        m.set_pattern(p)
        mu = m.run(3)
    assert r.stdout == expected_text(mu)


def test_cli_full_precision_bin_against_oracle(tmp_path):
    """The reference's own input format at full precision: FP64 features in the .bin files (the georgiy sample read as
    prob2bin.py reads it, data edges drawn in FP64), run through the CLI and through ac_set_graph64, against the FP64
    oracle on the same unrounded arrays.  The device stores features as FP32, so this measures that rounding too.  The
    CLI's binary MU dump and its top-k lines (most likely first, ties by lower vertex) are checked on the way."""
    c = cases.georgiy64()
    assert c.data.vfeat.dtype == np.float64 and (c.data.vfeat != c.data.vfeat.astype(np.float32)).any()
    db, pb = str(tmp_path / "data.bin"), str(tmp_path / "pattern.bin")
    save_bin64(db, c.data)
    save_bin64(pb, c.pattern)
    mub, topk = str(tmp_path / "mu.bin"), str(tmp_path / "topk.txt")
    r = run_cli([db, pb, "--mu-bin", mub, "--topk", "3", topk])
    assert r.returncode == 0, r.stderr
    raw = open(mub, "rb").read()
    dn, pn = (int(x) for x in np.frombuffer(raw[:16], np.uint64))
    assert (dn, pn) == (c.data.num_nodes, c.pattern.num_nodes)
    mu_cli = np.frombuffer(raw[16:], np.float32).reshape(dn, pn)
    assert r.stdout == expected_text(mu_cli)
    lines = [ln.split() for ln in open(topk).read().splitlines()]
    assert len(lines) == pn * 3
    for p in range(pn):
        order = sorted(range(dn), key=lambda v: (-float(mu_cli[v, p]), v))[:3]
        for k in range(3):
            pp, rank, v, x = lines[p * 3 + k]
            assert (int(pp), int(rank), int(v)) == (p, k, order[k]) and x == "%e" % float(mu_cli[order[k], p])
    with Matcher(0) as m:
        m.set_graph64(AC_DATA, c.data)
        m.set_graph64(AC_PATTERN, c.pattern)
        mu = m.run()
    np.testing.assert_array_equal(mu, mu_cli)
    ref = O.run(O.Graph64.of(c.data.vfeat, c.data.efeat, c.data.src, c.data.dst),
                O.Graph64.of(c.pattern.vfeat, c.pattern.efeat, c.pattern.src, c.pattern.dst))
    bad, ratio = allclose_report(mu, ref)
    print(f"georgiy64 (FP64 .bin) vs FP64 oracle on the unrounded arrays: max err/tol {ratio:.4f}", flush=True)
    assert bad == 0


def test_library_topk_after_set_graph64():
    """ac_get_topk through the library after ac_set_graph64 (whose sizes the Python Matcher records like
    set_data / set_pattern), against a sort of the same MU."""
    c = cases.georgiy64()
    with Matcher(0) as m:
        m.set_graph64(AC_DATA, c.data)
        m.set_graph64(AC_PATTERN, c.pattern)
        mu = m.run()
        idx, val = m.topk(3)
    for p in range(mu.shape[1]):
        order = sorted(range(mu.shape[0]), key=lambda v: (-float(mu[v, p]), v))[:3]
        assert list(idx[p]) == order
        np.testing.assert_array_equal(val[p], mu[order, p])
