"""Algorithmic FP64 operation count per primitive tuple of the benchmark
workload, from the host build of the integral engine with its operation
counter (tests/engine_host.cpp compiled with -DNRX_COUNT_FLOPS; FMA = 2).

    g++ -O2 -fPIC -shared -DNRX_COUNT_FLOPS -ffp-contract=off \
        -o tests/_build/libengine_flops.so tests/engine_host.cpp
    python tests/count_flops.py

The printed averages are the per-tuple figures bench.py multiplies by the
tuples of one launch for the roofline.
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "neo-rxchf_amd"))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from neorx import synth  # noqa: E402


def main():
    E = C.CDLL(os.path.join(ROOT, "tests", "_build", "libengine_flops.so"))
    E.eng_flops_per_tuple.restype = C.c_double
    s = synth.config2()
    P = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)  # noqa: E731
    F = lambda a: np.asfortranarray(a).ctypes.data_as(C.c_void_p)  # noqa: E731
    for cls, n, seed in [(1, None, None), (2, 20000, 2), (3, 5000, 3)]:
        if cls == 1:  # every 47th tuple of the full e-p enumeration
            ne, npb = s["npebf"], s["npbf"]
            r = np.arange(0, ne * ne * npb * npb, 47)
            idx = np.stack([(r // ne) % ne, r // ne // ne // npb, r % ne, (r // ne // ne) % npb], 1).astype(np.int64)
        else:
            idx = bench.sample_tuples(np, s, cls, n, seed)
        f = E.eng_flops_per_tuple(C.c_int64(s["npebf"]), C.c_int64(s["npbf"]), P(s["elcex"]), P(s["nucex"]),
                                  F(s["elcam"]), F(s["nucam"]), F(s["elcbfc"]), F(s["nucbfc"]), C.c_int64(s["nat"]),
                                  C.c_int64(s["ngtg"]), C.c_double(s["pmass"]), P(s["cat"]), P(s["zan"]),
                                  P(s["bcoef"]), P(s["gamma"]), C.c_int64(cls), C.c_int64(len(idx)), P(idx))
        print(f"class {cls}: {len(idx)} tuples, {f:.1f} flops/tuple")


if __name__ == "__main__":
    main()
