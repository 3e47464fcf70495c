"""Shared helpers of the test suite."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# Parity bar for FP64 integrals: |x - ref| <= RTOL |ref| + ATOL_REL * max|ref|.
# RTOL is the per-value relative bar; ATOL_REL bounds values that are
# themselves the residue of cancellation (below 1e-12 of the batch scale both
# implementations carry only rounding noise there).
RTOL = 1e-11
ATOL_REL = 1e-12


def check_close(got, ref, rtol=RTOL, atol_rel=ATOL_REL, what=""):
    got = np.asarray(got, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    assert got.shape == ref.shape, (got.shape, ref.shape)
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    err = np.abs(got - ref)
    bad = err > rtol * np.abs(ref) + atol_rel * scale
    if np.any(bad):
        i = int(np.argmax(err - rtol * np.abs(ref)))
        raise AssertionError(f"{what}: {int(bad.sum())}/{ref.size} values outside bar; worst idx {i}: "
                             f"got {got[i]!r} ref {ref[i]!r} scale {scale:.3e}")
    rel = err / np.maximum(np.abs(ref), 1e-300)
    return dict(max_abs_over_scale=float(err.max() / scale) if scale else 0.0,
                frac_rel_1e12=float(np.mean(rel <= 1e-12)))


def random_tuples(sysd, cls, n, seed):
    rng = np.random.default_rng(seed)
    w = 2 * (cls + 1)
    idx = np.zeros((n, w), dtype=np.int64)
    for i in list(range(cls)) + list(range(cls + 1, 2 * cls + 1)):
        idx[:, i] = rng.integers(0, sysd["npebf"], n)
    idx[:, cls] = rng.integers(0, sysd["npbf"], n)
    idx[:, 2 * cls + 1] = rng.integers(0, sysd["npbf"], n)
    return idx


_eng = {}


def engine_host(lmax=2):
    """Host build of the device integral engine (tests/_build) at the
    capacities of l <= 2 (the tier kernels) or l <= 3 (the wide bin)."""
    if lmax not in _eng:
        name = "libengine_host.so" if lmax == 2 else "libengine_host3.so"
        _eng[lmax] = C.CDLL(os.path.join(ROOT, "tests", "_build", name))
        _eng[lmax].eng_gint.restype = C.c_double
    return _eng[lmax]


def engine_prim_list(s, cls, idx, tier=0, lmax=2):
    """Host engine values; tier 0 picks the capacity as the device does,
    tier 1 forces the general capacity."""
    e, sv, _ = engine_prim_list3(s, cls, idx, tier, lmax)
    return e, sv


def mult_tuples(sysd, cls, n, seed):
    """Random primitive tuples of the RXCHFmult classes over the combined
    primitive list of a two-basis system: electron 1 from the regular
    primitives, the others from the special-electron ones."""
    rng = np.random.default_rng(seed)
    ne = 2 if cls == 5 else 3
    pa, pb, npb = sysd["npebf"], sysd["npebfBE"], sysd["npbf"]
    idx = np.zeros((n, 2 * (ne + 1)), dtype=np.int64)
    for off in (0, ne + 1):
        idx[:, off] = rng.integers(0, pa, n)
        for i in range(1, ne):
            idx[:, off + i] = pa + rng.integers(0, pb, n)
        idx[:, off + ne] = rng.integers(0, npb, n)
    return idx


def engine_prim_list3(s, cls, idx, tier=0, lmax=2):
    """As engine_prim_list with the third output (ansE2 of class 5)."""
    E = engine_host(lmax)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    n = len(idx)
    e, sv, xv = np.zeros(n), np.zeros(n), np.zeros(n)
    keep = [np.ascontiguousarray(s["elcex"]), np.ascontiguousarray(s["nucex"]), np.asfortranarray(s["elcam"]),
            np.asfortranarray(s["nucam"]), np.asfortranarray(s["elcbfc"]), np.asfortranarray(s["nucbfc"])]
    sk = [np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma")]
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    E.eng_prim_list(C.c_int64(s["npebf"]), C.c_int64(s["npbf"]), *[P(a) for a in keep], C.c_int64(s["nat"]),
                    C.c_int64(s["ngtg"]), C.c_double(s["pmass"]), *[P(a) for a in sk], C.c_int64(cls), C.c_int64(n),
                    P(idx), P(e), P(sv), P(xv), C.c_int64(tier))
    return e, sv, xv


def golden_scf():
    """tests/golden/xchf_scf.json: reference xcscf runs (make_xchf_scf.py)."""
    import json
    return json.load(open(os.path.join(ROOT, "tests", "golden", "xchf_scf.json")))


def golden_system(entry):
    """System dict of a golden entry: the stored arrays, or the reference
    input file it names (with its guess orbitals) -> (sysd, guess or None)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "neo-rxchf_amd"))
    from neorx import xchf
    if "input" in entry:
        gd = os.path.join(ROOT, "tests", "golden")
        s = xchf.read_input(os.path.join(gd, entry["input"]))
        guess = xchf.read_mos(os.path.join(gd, entry["guess"]), s["nebf"]) if entry.get("guess") else None
        return s, guess
    s = {}
    for k, v in entry["system"].items():
        if isinstance(v, dict) and "data" in v:
            a = np.array(v["data"], dtype=v["dtype"])
            s[k] = a.reshape(v["shape"], order=v["order"]) if v["shape"] else a
            if v["order"] == "F" and len(v["shape"]) > 1:
                s[k] = np.asfortranarray(s[k])
        else:
            s[k] = v
    return s, None
