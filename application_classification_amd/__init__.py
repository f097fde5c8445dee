"""MI355X-native ApplicationClassification matcher (data graph x pattern graph -> per-vertex log-probabilities).

The compute path is libac_mi355x.so (hand-written HIP kernels for gfx950 behind the C ABI in include/ac_mi355x.h);
this package is its host-side mirror of the reference driver and operator library.
"""
from ._lib import ACError  # noqa: F401
from .matcher import Matcher, ac, load_bin, load_csv, run_matcher, save_bin, save_csv  # noqa: F401
