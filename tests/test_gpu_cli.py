"""The command-line driver (ac_main, the reference's `main data.bin pattern.bin > result`, src/main.cu:66-257): its
stdout is one "%e" line per MU value in dn x pn order, byte for byte the formatting of the library's own result, and
stderr carries the `elapsed=` line (src/main.cu:250)."""
import os
import subprocess

import numpy as np
import pytest

from application_classification_amd import workloads as W
from application_classification_amd.matcher import Matcher, load_csv, save_bin, save_csv

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
