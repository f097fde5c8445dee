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


# --- the reference's Graph at full precision (ac_graph64) and prob2bin.py's conversion -----------------------------

def prob2bin_bytes(vertex_csv, edge_csv):
    """The bytes prob2bin.py writes for one CSV pair, by its own steps with pandas (prob2bin.py:11-12 read, :28-29
    id check, :33-39 column split, :46-64 layout) -- pandas is the independent parser the C++ reader is checked
    against."""
    import pandas as pd

    def read_csv(p):
        return pd.read_csv(p, skiprows=1, sep=" ", header=None)

    v, e = read_csv(vertex_csv), read_csv(edge_csv)
    assert (v[0] == v.index).all()
    vt = v.values[:, 1:]
    et = e[list(range(2, e.shape[1]))].values
    ee = e[[0, 1]].values
    parts = [np.hstack([vt.shape, et.shape]).astype(np.uint64), vt.ravel().astype(np.float64),
             et.ravel().astype(np.float64), ee[:, 0].ravel().astype(np.uint64), ee[:, 1].ravel().astype(np.uint64)]
    return b"".join(bytes(bytearray(x)) for x in parts)


def _write_edge_csv(path, n_vertices, n_edges, dim, seed):
    """A data-edge CSV in the reference's text form (header "N C", CRLF, values as Python's shortest repr), with
    magnitudes from 1e-30 to 1e30, negatives, integers and exponent forms."""
    r = np.random.default_rng(seed)
    src = r.integers(0, n_vertices, n_edges)
    dst = r.integers(0, n_vertices, n_edges)
    vals = r.normal(size=(n_edges, dim)) * 10.0 ** r.integers(-30, 30, size=(n_edges, dim))
    vals[:, 0] = r.uniform(0, 10, n_edges)
    vals[:, 1] = np.round(vals[:, 1] * 1e3)  # integer-valued floats
    lines = [f"{n_edges} {dim + 2}"]
    for i in range(n_edges):
        lines.append(" ".join([str(src[i]), str(dst[i])] + [repr(float(x)) for x in vals[i]]))
    open(path, "w", newline="").write("\r\n".join(lines) + "\r\n")


def test_csv_to_bin_is_byte_identical_to_prob2bin(tmp_path):
    exe = os.path.join(ROOT, "application_classification_amd", "ac_main")
    ref = os.path.join(cases.REF_DATA, "georgiyPattern")
    e_csv = str(tmp_path / "dataE.csv")
    _write_edge_csv(e_csv, 1000, 3000, 17, seed=3)
    pairs = [(ref + ".Vertex.csv", ref + ".Edges.csv"),
             (os.path.join(cases.REF_DATA, "georgiyData.Vertex.csv"), e_csv)]
    for i, (v, e) in enumerate(pairs):
        expect = prob2bin_bytes(v, e)
        out = str(tmp_path / f"g{i}.bin")
        matcher.save_bin64(out, matcher.load_csv64(v, e))
        assert open(out, "rb").read() == expect
        out2 = str(tmp_path / f"cli{i}.bin")
        r = subprocess.run([exe, "--to-bin", v, e, out2], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert open(out2, "rb").read() == expect
        # pandas' converter is not the correctly rounded one; this file differs from strtod somewhere
        toks = [float(t) for line in open(v).read().splitlines()[1:] for t in line.split()]
        g = matcher.load_csv64(v, e)
        assert (np.array(toks).reshape(g.num_nodes, -1)[:, 1:] != g.vfeat).any() or i == 1


def test_bin64_round_trip_and_validation(tmp_path):
    ref = os.path.join(cases.REF_DATA, "georgiyPattern")
    g = matcher.load_csv64(ref + ".Vertex.csv", ref + ".Edges.csv")
    assert g.vfeat.dtype == np.float64 and g.src.dtype == np.uint64
    path = str(tmp_path / "p.bin")
    matcher.save_bin64(path, g)
    h = matcher.load_bin64(path)
    for a, b in ((g.vfeat, h.vfeat), (g.efeat, h.efeat), (g.src, h.src), (g.dst, h.dst)):
        np.testing.assert_array_equal(a, b)
    # the FP32 reader is the FP64 one rounded
    f = matcher.load_bin(path)
    np.testing.assert_array_equal(f.vfeat, g.vfeat.astype(np.float32))
    np.testing.assert_array_equal(f.efeat, g.efeat.astype(np.float32))
    c = matcher.load_csv(ref + ".Vertex.csv", ref + ".Edges.csv")
    np.testing.assert_array_equal(c.efeat, g.efeat.astype(np.float32))
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-1])
    with pytest.raises(_lib.ACError) as e:
        matcher.load_bin64(path)
    assert e.value.code == _lib.AC_ERR_IO
    bad = bytearray(raw)
    bad[-8:] = np.array([10], np.uint64).tobytes()  # last destination = num_nodes
    open(path, "wb").write(bytes(bad))
    with pytest.raises(_lib.ACError) as e:
        matcher.load_bin64(path)
    assert e.value.code == _lib.AC_ERR_IO


@pytest.mark.parametrize("tok", ["0.1", "-2.5e-3", "1E5", "+3.25", ".5", "5.", "123456789012345678901234",
                                 "0.000000000000000000001234567890123456789", "1e-320", "-0.0", "7.0e+02",
                                 "9.999999999999999e22", "2.2250738585072014e-308"])
def test_decimal_conversion_matches_pandas(tmp_path, tok):
    import pandas as pd

    v, e = str(tmp_path / "v.csv"), str(tmp_path / "e.csv")
    open(v, "w").write(f"1 2\n0 {tok}\n")
    open(e, "w").write("0 2\n")
    g = matcher.load_csv64(v, e)
    want = pd.read_csv(v, skiprows=1, sep=" ", header=None).values[0, 1]
    assert np.float64(want).tobytes() == np.float64(g.vfeat[0, 0]).tobytes(), (tok, want, g.vfeat[0, 0])


def test_distance_square_root_correction_is_correctly_rounded(tmp_path):
    """The distance kernels' FP32 square root (the hardware root corrected by ac::sqrt_rn_fix) is correctly rounded:
    tests/sqrt_rn_check.hip runs the same correction step on the host over every 101st float from 2^-104 up, from
    the correctly rounded root and both its neighbours."""
    import os
    exe = str(tmp_path / "sqrt_rn_check")
    src = os.path.join(os.path.dirname(__file__), "sqrt_rn_check.hip")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-x", "hip", "--offload-arch=gfx950", src, "-o", exe],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "101"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and " 0 wrong" in r.stdout, r.stdout + r.stderr
