"""Golden vectors for primitive tuples with f functions (l = 3): the
reference's own tuple routines (xcalc_GAM1_MD, source/gam_1_OMP.f:398;
xcalc_GAM2_MD, source/gam_2_OMP.f:546; xcalc_GAM3_MD, source/gam_3_OMP.f:617),
built from the reference's sources into oracle/_ref/libref.so by
``make -C oracle ref``, evaluated in this container on two seeded systems:

* f_elec: s/p/f electronic shells on two centres, s/p/d nuclear shells;
* f_nuc:  s/p electronic shells, an s and an f nuclear shell.

Tuples are drawn at random among those with at least one f primitive.  The
reference's Fortran takes tens of seconds per e-e-p and e-e-e-p tuple with f
functions, so the CPU tests check the oracle against these stored values
instead of calling it.

Writes tests/golden/f_tuples.json: per system the input arrays and per class
the index tuples and the reference's ansE (and ansS for e-p and e-e-p).

    python tests/golden/make_f_tuples.py
"""
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "neo-rxchf_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)
import oracle  # noqa: E402
from neorx import synth  # noqa: E402
from make_xchf_scf import jsonable  # noqa: E402

SYSTEMS = {
    "f_elec": dict(seed=91, n_e_centres=2, e_shells=(0, 1, 3), contract=(1, 2), p_shells=(0, 1, 2), ngtg=3),
    "f_nuc": dict(seed=92, n_e_centres=2, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 3), ngtg=2),
}
COUNTS = {1: 24, 2: 6, 3: 3}


def f_tuples(s, cls, n, seed):
    """n random tuples of class cls with at least one f primitive."""
    rng = np.random.default_rng(seed)
    fe = np.asarray(s["elcam"]).max(axis=1) >= 3
    fp = np.asarray(s["nucam"]).max(axis=1) >= 3
    out = []
    while len(out) < n:
        t = np.zeros(2 * (cls + 1), dtype=np.int64)
        for i in list(range(cls)) + list(range(cls + 1, 2 * cls + 1)):
            t[i] = rng.integers(0, s["npebf"])
        t[cls] = rng.integers(0, s["npbf"])
        t[2 * cls + 1] = rng.integers(0, s["npbf"])
        el = [t[i] for i in list(range(cls)) + list(range(cls + 1, 2 * cls + 1))]
        if any(fe[el]) or fp[t[cls]] or fp[t[2 * cls + 1]]:
            out.append(t)
    return np.array(out)


def main():
    res = {}
    for name, kw in SYSTEMS.items():
        s = synth.small(**kw)
        ent = {"system": jsonable(s), "tuples": {}}
        for cls, n in COUNTS.items():
            idx = f_tuples(s, cls, n, seed=100 * cls + len(name))
            with ThreadPoolExecutor(max_workers=8) as ex:
                vals = list(ex.map(lambda r: oracle.ref_prim_list(s, cls, r[None, :]), idx))
            E = [float(v[0][0]) for v in vals]
            S = [float(v[1][0]) for v in vals] if cls < 3 else None
            ent["tuples"][str(cls)] = {"idx": idx.tolist(), "E": E, "S": S}
            print(name, cls, "done", flush=True)
        res[name] = ent
    with open(os.path.join(HERE, "f_tuples.json"), "w") as f:
        json.dump(res, f, indent=0)


if __name__ == "__main__":
    main()
