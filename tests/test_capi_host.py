"""Host-side checks of libac_mi355x.so that need no GPU: exports, generators, file formats, the CLI driver."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from application_classification_amd import _lib, matcher
from application_classification_amd import workloads as W
from tests.conftest import ROOT
from tests import cases

HEADER = os.path.join(ROOT, "include", "ac_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ac_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures and the header disagree"
    _lib.lib()  # declares every signature


def test_create_without_device_reports_nodev():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = _lib.lib().ac_create(ctypes.byref(h), 0)
    assert rc == _lib.AC_ERR_NODEV
    assert b"device" in _lib.lib().ac_global_error()


def test_er_generator():
    s1, d1 = W.er_edges(100, 900, 5)
    s2, d2 = W.er_edges(100, 900, 5)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_array_equal(d1, d2)
    keys = s1.astype(np.int64) * 100 + d1
    assert len(np.unique(keys)) == 900 and (s1 != d1).all() and (np.diff(keys) > 0).all()
    s3, _ = W.er_edges(100, 900, 6)
    assert not np.array_equal(s1, s3)


def test_rmat_generator():
    s, d = W.rmat_edges(12, 16, seed=0)
    n = 1 << 12
    keys = s.astype(np.int64) * n + d
    assert (s != d).all() and (np.diff(keys) > 0).all()  # no self-loops, sorted, no duplicates
    assert 0.8 * 16 * n < len(s) <= 16 * n
    outdeg = np.bincount(s, minlength=n)
    assert outdeg.max() > 20 * outdeg.mean()  # skewed: a = 0.57 concentrates edges on low ids
    s2, d2 = W.rmat_edges(12, 16, seed=0)
    np.testing.assert_array_equal(s, s2)


def test_normal_generator():
    x = W.normal((200000,), 3, 10)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    np.testing.assert_array_equal(x, W.normal((200000,), 3, 10))
    assert not np.array_equal(x, W.normal((200000,), 3, 11))


def test_planted_pattern_is_a_connected_induced_subgraph():
    w = W.small(0)
    p = w.pattern
    assert len(set(p.vmap.tolist())) == 10
    np.testing.assert_array_equal(w.data.src[p.emap], p.vmap[p.src])
    np.testing.assert_array_equal(w.data.dst[p.emap], p.vmap[p.dst])
    # connected (undirected)
    seen = {0}
    changed = True
    while changed:
        changed = False
        for a, b in zip(p.src, p.dst):
            if (a in seen) != (b in seen):
                seen |= {a, b}
                changed = True
    assert seen == set(range(10))
    noise = p.vfeat - w.data.vfeat[p.vmap]
    assert 0.005 < noise.std() < 0.02


def test_bin_round_trip(tmp_path):
    c = cases.tiny_hand()
    path = str(tmp_path / "g.bin")
    matcher.save_bin(path, c.data)
    g = matcher.load_bin(path)
    np.testing.assert_array_equal(g.vfeat, c.data.vfeat)
    np.testing.assert_array_equal(g.efeat, c.data.efeat)
    np.testing.assert_array_equal(g.src, c.data.src)
    np.testing.assert_array_equal(g.dst, c.data.dst)
    # the layout of prob2bin.py:46-64
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:32], np.uint64)
    np.testing.assert_array_equal(hdr, [6, 4, 9, 5])
    assert len(raw) == 32 + 8 * (6 * 4 + 9 * 5 + 2 * 9)
    np.testing.assert_array_equal(np.frombuffer(raw[32:32 + 8 * 24], np.float64).reshape(6, 4), c.data.vfeat)


def test_bin_rejects_truncated_and_bad_ids(tmp_path):
    c = cases.tiny_hand()
    path = str(tmp_path / "g.bin")
    matcher.save_bin(path, c.data)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-8])
    with pytest.raises(_lib.ACError) as e:
        matcher.load_bin(path)
    assert e.value.code == _lib.AC_ERR_IO
    bad = bytearray(raw)
    off = len(raw) - 8  # last destination id
    bad[off:off + 8] = np.array([6], np.uint64).tobytes()
    open(path, "wb").write(bytes(bad))
    with pytest.raises(_lib.ACError):
        matcher.load_bin(path)
    with pytest.raises(_lib.ACError):
        matcher.load_bin(str(tmp_path / "missing.bin"))


def test_csv_round_trip_and_reference_files(tmp_path):
    c = cases.tiny_hand()
    v, e = str(tmp_path / "v.csv"), str(tmp_path / "e.csv")
    matcher.save_csv(v, e, c.data)
    g = matcher.load_csv(v, e)
    np.testing.assert_array_equal(g.vfeat, c.data.vfeat)
    np.testing.assert_array_equal(g.efeat, c.data.efeat)
    np.testing.assert_array_equal(g.src, c.data.src)
    # the reference's own georgiy pattern files (CRLF, header "N C")
    ref = os.path.join(cases.REF_DATA, "georgiyPattern")
    g = matcher.load_csv(ref + ".Vertex.csv", ref + ".Edges.csv")
    assert g.vfeat.shape == (10, 13) and g.efeat.shape == (21, 17)
    raw = np.loadtxt(ref + ".Edges.csv", skiprows=1)
    np.testing.assert_array_equal(g.src, raw[:, 0].astype(np.int32))
    np.testing.assert_array_equal(g.efeat, raw[:, 2:].astype(np.float32))


def test_csv_rejects_bad_rows(tmp_path):
    v, e = str(tmp_path / "v.csv"), str(tmp_path / "e.csv")
    open(v, "w").write("2 2\n0 1.5\n2 2.5\n")  # second id is not the row index
    open(e, "w").write("1 3\n0 1 0.5\n")
    with pytest.raises(_lib.ACError):
        matcher.load_csv(v, e)
    open(v, "w").write("2 2\n0 1.5\n1 2.5\n")
    open(e, "w").write("1 3\n0 7 0.5\n")  # destination out of range
    with pytest.raises(_lib.ACError):
        matcher.load_csv(v, e)
    open(e, "w").write("2 3\n0 1 0.5\n1 0 0.5 9\n")  # ragged
    with pytest.raises(_lib.ACError):
        matcher.load_csv(v, e)


def test_cli_usage():
    exe = os.path.join(ROOT, "application_classification_amd", "ac_main")
    assert os.access(exe, os.X_OK)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
