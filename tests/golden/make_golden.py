"""Writes the golden fixtures of tests/golden/ from the FP64 oracle (oracle/ac_oracle.cpp).

The reference holds no usable expected outputs (its results/ and rmat inputs are absent), so the expected values
here come from the oracle, which tests/test_oracle_pins.py pins independently.  Inputs are regenerated from seeds
by application_classification_amd.workloads; each fixture stores a checksum of its inputs so that a change of the
generator is reported as such rather than as a parity failure.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from application_classification_amd import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        h.update(str(a.dtype).encode() + str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def write(name, data, patt, n_iter=-1):
    D = O.Graph64.of(data.vfeat, data.efeat, data.src, data.dst)
    P = O.Graph64.of(patt.vfeat, patt.efeat, patt.src, patt.dst)
    mu = O.run(D, P, n_iter=n_iter, threads=1)
    key = digest(data.vfeat, data.efeat, data.src, data.dst, patt.vfeat, patt.efeat, patt.src, patt.dst)
    np.savez(os.path.join(HERE, f"{name}.npz"), mu=mu, inputs_sha256=np.array(key), n_iter=np.array(n_iter))
    print(name, mu.shape, key[:16])


def main():
    w = W.small(0)
    write("small_er1000", w.data, w.pattern)
    g = cases.georgiy()
    write("georgiy", g.data, g.pattern)
    h = cases.tiny_hand()
    write("tiny_hand", h.data, h.pattern)


if __name__ == "__main__":
    main()
