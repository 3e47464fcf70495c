"""ctypes bindings of the CPU checkers (TEST INFRASTRUCTURE ONLY).

* ``liboracle.so``  -- this repository's C restatement (neorx_oracle.c);
* ``_ref/libref.so`` -- the reference's own Fortran hot-path sources compiled
  by ``make -C oracle ref`` (present only where /root/reference was mounted at
  build time).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_P = C.c_void_p
_o = None
_r = None


def _p(a):
    return a.ctypes.data_as(_P)


def lib():
    global _o
    if _o is None:
        _o = C.CDLL(ORACLE_SO)
        _o.orc_boys.restype = C.c_double
        _o.orc_gint.restype = C.c_double
    return _o


def _bas(s):
    keep = [np.ascontiguousarray(s[k]) for k in ("kpestr", "kpeend", "ampeb2c", "agebfcc", "agnbfcc", "elcex",
                                                 "nucex")]
    keep += [np.asfortranarray(s[k]) for k in ("elcam", "nucam", "elcbfc", "nucbfc")]
    args = [C.c_int64(s["nebf"]), C.c_int64(s["npebf"]), C.c_int64(s["npbf"])] + [_p(a) for a in keep]
    return args, keep


def _sysargs(s):
    keep = [np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma")]
    args = [C.c_int64(s["nat"]), C.c_int64(s["ngtg"]), C.c_double(s["pmass"])] + [_p(a) for a in keep]
    return args, keep


def boys(m, T):
    return lib().orc_boys(C.c_int64(m), C.c_double(T))


def prim_list(s, cls, idx):
    """Primitive tuple values (xcalc_GAM{1,2,3}_MD semantics) for an
    (ntup, 2*(cls+1)) int64 array of 0-based primitive indices."""
    e, sv, _ = prim_list3(s, cls, idx)
    return e, sv


def prim_list3(s, cls, idx):
    """As prim_list, with the third output (ansE2 of the RXCHFmult e-e-p class
    5).  For classes 5 and 6 s is a one-basis dict over the combined primitive
    list (synth.combined)."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    n = idx.shape[0]
    e, sv, e2 = np.zeros(n), np.zeros(n), np.zeros(n)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_prim_list3(*ba, *sa, C.c_int64(cls), C.c_int64(n), _p(idx), _p(e), _p(sv), _p(e2))
    return e, sv, e2


def gam1(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne * ne * npb * npb)
    gs = np.zeros_like(g)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_gam1(*ba, *sa, _p(g), _p(gs))
    return g, gs


def gam2_ic1(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne ** 4 * npb * npb)
    gs = np.zeros_like(g)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_gam2_ic1(*ba, *sa, _p(g), _p(gs))
    return g, gs


def gam3_ic1(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne ** 6 * npb * npb)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_gam3_ic1(*ba, *sa, _p(g))
    return g


def gam_ee(s):
    ne = s["nebf"]
    g = np.zeros(ne ** 4)
    keep = [np.ascontiguousarray(s["ampeb2c"]), np.ascontiguousarray(s["agebfcc"]), np.ascontiguousarray(s["elcex"]),
            np.asfortranarray(s["elcam"]), np.asfortranarray(s["elcbfc"])]
    lib().orc_gam_ee(C.c_int64(ne), C.c_int64(s["npebf"]), *[_p(a) for a in keep], _p(g))
    return g


def gam4_icr(nebf, npbf, g2s, gee):
    g = np.zeros(nebf ** 8 * npbf * npbf)
    lib().orc_gam4_icr(C.c_int64(nebf), C.c_int64(npbf), _p(np.ascontiguousarray(g2s)),
                       _p(np.ascontiguousarray(gee)), _p(g))
    return g


def eg2ic1(nebf, npbf, DE, DP, G, Gs, FE, FP):
    DE = np.asfortranarray(DE)
    DP = np.asfortranarray(DP)
    SE = np.zeros((nebf, nebf), order="F")
    SP = np.zeros((npbf, npbf), order="F")
    e = np.zeros(1)
    s = np.zeros(1)
    lib().orc_eg2ic1(C.c_int64(nebf), C.c_int64(npbf), _p(DE), _p(DP), _p(np.ascontiguousarray(G)),
                     _p(np.ascontiguousarray(Gs)), _p(FE), _p(SE), _p(FP), _p(SP), _p(e), _p(s))
    return SE, SP, float(e[0]), float(s[0])


def eg3ic1(nebf, npbf, DE, DP, G, FE, FP):
    e = np.zeros(1)
    lib().orc_eg3ic1(C.c_int64(nebf), C.c_int64(npbf), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)),
                     _p(np.ascontiguousarray(G)), _p(FE), _p(FP), _p(e))
    return float(e[0])


def eg4icr(nebf, npbf, DE, DP, G, FE, FP):
    e = np.zeros(1)
    lib().orc_eg4icr(C.c_int64(nebf), C.c_int64(npbf), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)),
                     _p(np.ascontiguousarray(G)), _p(FE), _p(FP), _p(e))
    return float(e[0])


def gam2_dscf(s, DE, DP, FE, FP, SE2, SP2):
    """GAM2_DSCF semantics: FE, FP, SE2, SP2 (Fortran-ordered arrays) are
    accumulated in place; returns (E_gam2, S_gam2)."""
    e, sv = np.zeros(1), np.zeros(1)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_gam2_dscf(*ba, *sa, _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE), _p(FP), _p(SE2),
                        _p(SP2), _p(e), _p(sv))
    return float(e[0]), float(sv[0])


def gam3_dscf(s, DE, DP, FE, FP):
    e = np.zeros(1)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_gam3_dscf(*ba, *sa, _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE), _p(FP), _p(e))
    return float(e[0])


def gam4_dscf(nebf, npbf, g2s, gee, DE, DP, FE, FP):
    e = np.zeros(1)
    lib().orc_gam4_dscf(C.c_int64(nebf), C.c_int64(npbf), _p(np.ascontiguousarray(g2s)),
                        _p(np.ascontiguousarray(gee)), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE),
                        _p(FP), _p(e))
    return float(e[0])


# ---------------------------------------------------------------------------
# the reference itself (Fortran ABI: lower case + '_', every argument by
# reference, INTEGER*8, column-major arrays)
# ---------------------------------------------------------------------------
def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _r
    if _r is None:
        _r = C.CDLL(REF_SO)
    return _r


def _i8(v):
    return C.byref(C.c_int64(int(v)))


def _d(v):
    return C.byref(C.c_double(float(v)))


def ref_prim_list(s, cls, idx):
    """xcalc_GAM{1,2,3}_MD of the reference for each tuple (cls 4:
    RXCHFne_xcalc_GAM2_MD, two outputs returned as E and S)."""
    e, sv, _ = ref_prim_list3(s, cls, idx)
    return e, sv


def ref_prim_list3(s, cls, idx):
    """As ref_prim_list; cls 5: RXCHFmult_GAM2_MD (ansE1, ansS, ansE2),
    cls 6: RXCHFmult_INT_GAM3_MD, over a combined one-basis dict; cls 9:
    RXCHF_xcalc_GAM2_MD (ansE1, ansS, ansE2); cls 10: RXCHF_xcalc_GAM3_MD
    (ans1, ans2 as E and S)."""
    R = ref()
    fn = {1: R.xcalc_gam1_md_, 2: R.xcalc_gam2_md_, 3: R.xcalc_gam3_md_, 4: R.rxchfne_xcalc_gam2_md_,
          5: R.rxchfmult_gam2_md_, 6: R.rxchfmult_int_gam3_md_, 7: R.rxchfmult_xchf_gam3_md_,
          8: R.rxchfmult_xcalc_gam1_md_, 9: R.rxchf_xcalc_gam2_md_, 10: R.rxchf_xcalc_gam3_md_}[cls]
    nel = {4: 2, 5: 2, 6: 3, 7: 3, 8: 1, 9: 2, 10: 3}.get(cls, cls)
    am_e, am_p = np.asarray(s["elcam"]), np.asarray(s["nucam"])
    c_e, c_p = np.asarray(s["elcbfc"]), np.asarray(s["nucbfc"])
    cat = np.ascontiguousarray(s["cat"])
    zan, b, g = (np.ascontiguousarray(s[k]) for k in ("zan", "bcoef", "gamma"))
    out_e = np.zeros(len(idx))
    out_s = np.zeros(len(idx))
    out_x = np.zeros(len(idx))
    for t, x in enumerate(np.asarray(idx)):
        keep, args = [], []

        def prim(am, ex, cen, i):
            c = np.ascontiguousarray(cen[i], dtype=np.float64)
            keep.append(c)
            return [_i8(am[i, 0]), _i8(am[i, 1]), _i8(am[i, 2]), _d(ex[i]), _p(c)]

        for i in range(nel):
            args += prim(am_e, s["elcex"], c_e, x[i])
        args += prim(am_p, s["nucex"], c_p, x[nel])
        for i in range(nel):
            args += prim(am_e, s["elcex"], c_e, x[nel + 1 + i])
        args += prim(am_p, s["nucex"], c_p, x[2 * nel + 1])
        e, sv, xv = C.c_double(), C.c_double(), C.c_double()
        tail = [_i8(s["nat"]), _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(g), C.byref(e)]
        if cls in (5, 9):  # ansE1, ansE2, ansS
            tail += [C.byref(xv), C.byref(sv)]
        elif cls not in (3, 6, 7):
            tail.append(C.byref(sv))
        fn(*args, *tail)
        out_e[t] = e.value
        out_s[t] = sv.value
        out_x[t] = xv.value
    return out_e, out_s, out_x


def _ref_basis_args(s):
    keep = [np.ascontiguousarray(s["kpestr"]), np.ascontiguousarray(s["kpeend"]), np.ascontiguousarray(s["ampeb2c"]),
            np.ascontiguousarray(s["agebfcc"]), np.ascontiguousarray(s["agnbfcc"]), np.ascontiguousarray(s["elcex"]),
            np.ascontiguousarray(s["nucex"]), np.asfortranarray(s["elcam"]), np.asfortranarray(s["nucam"]),
            np.asfortranarray(s["elcbfc"]), np.asfortranarray(s["nucbfc"])]
    return keep


def ref_gam2_ic1(s):
    """GAM2_IC1 (source/gam2_IC1_OMP.f:2) of the reference."""
    ne, npb = s["nebf"], s["npbf"]
    ng2 = ne ** 4 * npb * npb
    g = np.zeros(ng2)
    gs = np.zeros(ng2)
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().gam2_ic1_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng2), _i8(0), _i8(s["nat"]), _i8(s["ngtg"]),
                    _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep], _p(g), _p(gs))
    return g, gs


def ref_gam3_ic1(s):
    ne, npb = s["nebf"], s["npbf"]
    ng3 = ne ** 6 * npb * npb
    g = np.zeros(ng3)
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().gam3_ic1_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng3), _i8(0), _i8(s["nat"]), _i8(s["ngtg"]),
                    _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep], _p(g))
    return g


def ref_eg2ic1(nebf, npbf, DE, DP, G, Gs, FE, FP):
    SE = np.zeros((nebf, nebf), order="F")
    SP = np.zeros((npbf, npbf), order="F")
    e = C.c_double()
    sv = C.c_double()
    ng2 = nebf ** 4 * npbf * npbf
    ref().eg2ic1_(_i8(1), _i8(nebf), _i8(npbf), _i8(ng2), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)),
                  _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gs)), _p(FE), _p(SE), _p(FP), _p(SP),
                  C.byref(e), C.byref(sv))
    return SE, SP, e.value, sv.value


def ref_eg3ic1(nebf, npbf, DE, DP, G, FE, FP):
    e = C.c_double()
    ng3 = nebf ** 6 * npbf * npbf
    ref().eg3ic1_(_i8(1), _i8(nebf), _i8(npbf), _i8(ng3), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)),
                  _p(np.ascontiguousarray(G)), _p(FE), _p(FP), C.byref(e))
    return e.value


def ref_eg4icr(nebf, npbf, DE, DP, G, FE, FP):
    e = C.c_double()
    ng4 = nebf ** 8 * npbf * npbf
    ref().eg4icr_(_i8(1), _i8(nebf), _i8(npbf), _i8(ng4), _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)),
                  _p(np.ascontiguousarray(G)), _p(FE), _p(FP), C.byref(e))
    return e.value


def ref_boys(m, T):
    F = np.zeros(m + 1)
    ref().xgammaf_(_p(F), _d(T), _i8(m))
    return F[m]


def ref_gam2_dscf(s, DE, DP, FE, FP, SE2, SP2):
    """GAM2_DSCF (source/gam2_DSCF_OMP.f:2) of the reference; arrays as in
    gam2_dscf."""
    ne, npb = s["nebf"], s["npbf"]
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    e, sv = C.c_double(), C.c_double()
    ref().gam2_dscf_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ne ** 4 * npb * npb), _i8(0), _i8(s["nat"]),
                     _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep],
                     _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE), _p(FP), _p(SE2), _p(SP2),
                     C.byref(e), C.byref(sv))
    return e.value, sv.value


def ref_gam3_dscf(s, DE, DP, FE, FP):
    ne, npb = s["nebf"], s["npbf"]
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    e = C.c_double()
    ref().gam3_dscf_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ne ** 6 * npb * npb), _i8(0), _i8(s["nat"]),
                     _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep],
                     _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE), _p(FP), C.byref(e))
    return e.value


def ref_gam4_dscf(nebf, npbf, g2s, gee, DE, DP, FE, FP, workdir):
    """GAM4_DSCF (source/gam4_DSCF_OMP.f:2) reads GAM_ee.ufm and GAM_2s.ufm
    (direct access, 8-byte records) from the working directory: they are
    written into workdir, which is made current for the call."""
    np.ascontiguousarray(gee, dtype=np.float64).tofile(os.path.join(workdir, "GAM_ee.ufm"))
    np.ascontiguousarray(g2s, dtype=np.float64).tofile(os.path.join(workdir, "GAM_2s.ufm"))
    ngee, ng2 = nebf ** 4, nebf ** 4 * npbf ** 2
    e = C.c_double()
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        ref().gam4_dscf_(_i8(1), _i8(nebf), _i8(npbf), _i8(ngee), _i8(ng2), _i8(ngee * ngee * npbf * npbf),
                         _p(np.asfortranarray(DE)), _p(np.asfortranarray(DP)), _p(FE), _p(FP), C.byref(e))
    finally:
        os.chdir(cwd)
    return e.value


def _in_dir(workdir, fn):
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        return fn()
    finally:
        os.chdir(cwd)


def ref_gam1(s, workdir):
    """GAM1_OMP_MD (source/gam_1_OMP.f:2) of the reference: it writes GAM_1.ufm
    and GAM_1s.ufm (direct access, 8-byte records) into workdir."""
    ne, npb = s["nebf"], s["npbf"]
    keep = [np.ascontiguousarray(s["ampeb2c"]), np.ascontiguousarray(s["agebfcc"]),
            np.ascontiguousarray(s["agnbfcc"]), np.ascontiguousarray(s["elcex"]), np.ascontiguousarray(s["nucex"]),
            np.asfortranarray(s["elcam"]), np.asfortranarray(s["nucam"]), np.asfortranarray(s["elcbfc"]),
            np.asfortranarray(s["nucbfc"])]
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    _in_dir(workdir, lambda: ref().gam1_omp_md_(_i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ne * ne * npb * npb),
                                                 _i8(s["npebf"] ** 2 * npb * npb),
                                                 _i8(s["nat"]), _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan),
                                                 _p(b), _p(gm), *[_p(a) for a in keep]))
    g = np.fromfile(os.path.join(workdir, "GAM_1.ufm"), dtype=np.float64)
    gs = np.fromfile(os.path.join(workdir, "GAM_1s.ufm"), dtype=np.float64)
    return g, gs


def ref_gam_ee(s, workdir):
    """calc_GAM_ee (source/gam_ee.f:2): writes GAM_ee.ufm into workdir."""
    ne = s["nebf"]
    keep = [np.ascontiguousarray(s["ampeb2c"]), np.ascontiguousarray(s["agebfcc"]), np.ascontiguousarray(s["elcex"]),
            np.asfortranarray(s["elcam"]), np.asfortranarray(s["elcbfc"])]
    _in_dir(workdir, lambda: ref().calc_gam_ee_(_i8(ne), _i8(s["npebf"]), _i8(ne ** 4), *[_p(a) for a in keep]))
    return np.fromfile(os.path.join(workdir, "GAM_ee.ufm"), dtype=np.float64)


def ref_gam4_icr(nebf, npbf, g2s, gee, workdir):
    """GAM4_ICR (source/gam4_ICR_OMP.f:2): reads GAM_ee.ufm from workdir."""
    np.ascontiguousarray(gee, dtype=np.float64).tofile(os.path.join(workdir, "GAM_ee.ufm"))
    ngee, ng2 = nebf ** 4, nebf ** 4 * npbf ** 2
    g4 = np.zeros(ngee * ngee * npbf * npbf)
    _in_dir(workdir, lambda: ref().gam4_icr_(_i8(1), _i8(nebf), _i8(npbf), _i8(ngee), _i8(ng2), _i8(g4.size),
                                             _p(np.ascontiguousarray(g2s)), _p(g4)))
    return g4


# ---------------------------------------------------------------------------
# XCUHF contractions (oracle and reference); F* arrays accumulate in place
# ---------------------------------------------------------------------------
def _f(a):
    return _p(np.asfortranarray(a))


def fae_omg2(nae, nbe, ne, npb, DAE, DBE, DP, G, Gs, FAE, SAE2):
    e, s = np.zeros(1), np.zeros(1)
    lib().orc_fae_omg2(C.c_int64(nae), C.c_int64(nbe), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                       _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gs)), _p(FAE), _p(SAE2), _p(e), _p(s))
    return float(e[0]), float(s[0])


def ufp_omg2(nae, nbe, ne, npb, DAE, DBE, DP, G, Gs, FP, SP2):
    e, s = np.zeros(1), np.zeros(1)
    lib().orc_ufp_omg2(C.c_int64(nae), C.c_int64(nbe), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                       _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gs)), _p(FP), _p(SP2), _p(e), _p(s))
    return float(e[0]), float(s[0])


def _omg(name, nae, nbe, ne, npb, DAE, DBE, DP, G, F):
    e = np.zeros(1)
    getattr(lib(), name)(C.c_int64(nae), C.c_int64(nbe), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                         _p(np.ascontiguousarray(G)), _p(F), _p(e))
    return float(e[0])


def fae_omg3(*a):
    return _omg("orc_fae_omg3", *a)


def ufp_omg3(*a):
    return _omg("orc_ufp_omg3", *a)


def fae_omg4(*a):
    return _omg("orc_fae_omg4", *a)


def ufp_omg4(*a):
    return _omg("orc_ufp_omg4", *a)


def omg4_icr(nebf, npbf, g2s, gee):
    g = np.zeros(nebf ** 8 * npbf * npbf)
    lib().orc_omg4_icr(C.c_int64(nebf), C.c_int64(npbf), _p(np.ascontiguousarray(g2s)), _p(np.ascontiguousarray(gee)),
                       _p(g))
    return g


def ref_fae_omg2(nae, nbe, ne, npb, DAE, DBE, DP, G, Gs, FAE, SAE2):
    """FAE_OMG2 (source/xcuhf/xcuhf_omg2.f:2) of the reference."""
    e, s = C.c_double(), C.c_double()
    DT = np.asfortranarray(DAE + DBE)
    ref().fae_omg2_(_i8(1), _i8(nae), _i8(nbe), _i8(ne), _i8(npb), _i8(ne ** 4 * npb * npb), _f(DAE), _f(DBE),
                    _p(DT), _f(DP), _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gs)), _p(FAE), _p(SAE2),
                    C.byref(e), C.byref(s))
    return e.value, s.value


def ref_ufp_omg2(nae, nbe, ne, npb, DAE, DBE, DP, G, Gs, FP, SP2):
    e, s = C.c_double(), C.c_double()
    DT = np.asfortranarray(DAE + DBE)
    ref().ufp_omg2_(_i8(1), _i8(ne), _i8(npb), _i8(ne ** 4 * npb * npb), _i8(nae), _i8(nbe), _f(DAE), _f(DBE),
                    _p(DT), _f(DP), _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gs)), _p(FP), _p(SP2),
                    C.byref(e), C.byref(s))
    return e.value, s.value


def _ref_omg(name, n, nae, nbe, ne, npb, DAE, DBE, DP, G, F):
    e = C.c_double()
    getattr(ref(), name)(_i8(1), _i8(ne), _i8(npb), _i8(ne ** (2 * n) * npb * npb), _i8(nae), _i8(nbe), _f(DAE),
                         _f(DBE), _f(DP), _p(np.ascontiguousarray(G)), _p(F), C.byref(e))
    return e.value


def ref_fae_omg3(*a):
    return _ref_omg("fae_omg3_", 3, *a)


def ref_ufp_omg3(*a):
    return _ref_omg("ufp_omg3_", 3, *a)


def ref_fae_omg4(*a):
    return _ref_omg("fae_omg4_", 4, *a)


def ref_ufp_omg4(*a):
    return _ref_omg("ufp_omg4_", 4, *a)


def ref_omg4_icr(nebf, npbf, g2s, gee, workdir):
    """OMG4_ICR (source/xcuhf/xcuhf_omg4.f:1080): reads GAM_ee.ufm from workdir."""
    np.ascontiguousarray(gee, dtype=np.float64).tofile(os.path.join(workdir, "GAM_ee.ufm"))
    ngee, ng2 = nebf ** 4, nebf ** 4 * npbf ** 2
    g4 = np.zeros(ngee * ngee * npbf * npbf)
    _in_dir(workdir, lambda: ref().omg4_icr_(_i8(1), _i8(nebf), _i8(npbf), _i8(ngee), _i8(ng2), _i8(g4.size),
                                             _p(np.ascontiguousarray(g2s)), _p(g4)))
    return g4


# ---------------------------------------------------------------------------
# RXCHF-ne (oracle and reference)
# ---------------------------------------------------------------------------
def rxchfne_gam2_ic1(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne ** 4 * npb * npb)
    gx = np.zeros_like(g)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfne_gam2_ic1(*ba, *sa, _p(g), _p(gx))
    return g, gx


def rxchfne_omg2(which, ne, npb, DAE, DBE, DP, G, F):
    """which: 0 RXCHFne_FAE_OMG2, 1 RXCHFne_FP_OMG2, 2 RXCHFne_FBE_OMG2."""
    e = np.zeros(1)
    lib().orc_rxchfne_omg2(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                           _p(np.ascontiguousarray(G)), _p(F), _p(e))
    return float(e[0])


def rxcuhfne_omg2(which, ne, npb, DAa, DAt, DBE, DP, G, Gx, F, F2=None):
    """which: 0 RXCUHFne_FE_OMG2 (F = FAalpE, F2 = FAbetE), 1 _FP_OMG2, 2 _FBE_OMG2."""
    e = np.zeros(1)
    lib().orc_rxcuhfne_omg2(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAa), _f(DAt), _f(DBE), _f(DP),
                            _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gx)), _p(F), _p(F2), _p(e))
    return float(e[0])


def rxchfne_fae_gamee(ne, gee, DAE, F):
    e = np.zeros(1)
    lib().orc_rxchfne_fae_gamee(C.c_int64(ne), _p(np.ascontiguousarray(gee)), _f(DAE), _p(F), _p(e))
    return float(e[0])


def ref_rxchfne_gam2_ic1(s):
    """RXCHFne_GAM2_IC1 (source/rxchfne/rxchfne_gam2_IC1_OMP.f:2)."""
    ne, npb = s["nebf"], s["npbf"]
    ng2 = ne ** 4 * npb * npb
    g, gx = np.zeros(ng2), np.zeros(ng2)
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfne_gam2_ic1_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng2), _i8(0), _i8(s["nat"]),
                            _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep],
                            _p(g), _p(gx))
    return g, gx


def ref_rxchfne_omg2(which, ne, npb, DAE, DBE, DP, G, F, nae=1, nbe=2):
    e = C.c_double()
    ng2 = ne ** 4 * npb * npb
    R = ref()
    Gp = _p(np.ascontiguousarray(G))
    if which == 0:
        R.rxchfne_fae_omg2_(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _f(DAE), _f(DBE), _f(DP), Gp, _p(F), C.byref(e))
    else:
        fn = R.rxchfne_fp_omg2_ if which == 1 else R.rxchfne_fbe_omg2_
        fn(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _i8(nae), _i8(nbe), _f(DAE), _f(DBE), _f(DP), Gp, _p(F), C.byref(e))
    return e.value


def ref_rxcuhfne_omg2(which, ne, npb, DAa, DAt, DBE, DP, G, Gx, F, F2=None, nab=1, nbe=1):
    e = C.c_double()
    ng2 = ne ** 4 * npb * npb
    R = ref()
    Gp, Gxp = _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(Gx))
    if which == 0:
        DAb = np.asfortranarray(np.asarray(DAt) - np.asarray(DAa))
        R.rxcuhfne_fe_omg2_(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _f(DAa), _p(DAb), _f(DAt), _f(DBE), _f(DP), Gp,
                            Gxp, _p(F), _p(F2), C.byref(e))
    else:
        fn = R.rxcuhfne_fp_omg2_ if which == 1 else R.rxcuhfne_fbe_omg2_
        fn(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _i8(nab), _i8(nbe), _f(DAa), _f(DAt), _f(DBE), _f(DP), Gp, Gxp,
           _p(F), C.byref(e))
    return e.value


def ref_rxchfne_fae_gamee(ne, gee, DAE, F):
    e = C.c_double()
    ref().rxchfne_fae_gamee_(_i8(ne), _i8(ne ** 4), _p(np.ascontiguousarray(gee)), _f(DAE), _p(F), C.byref(e))
    return e.value


# ---------------------------------------------------------------------------
# RXCHFmult (oracle and reference); s is a two-basis dict (synth.mult)
# ---------------------------------------------------------------------------
def _mbas(s):
    c, f = np.ascontiguousarray, np.asfortranarray
    keep = [c(s["kpestr"]), c(s["kpeend"]), c(s["agebfcc"]), c(s["elcex"]), f(s["elcam"]), f(s["elcbfc"]),
            c(s["kpestr_be"]), c(s["kpeend_be"]), c(s["agebfcc_be"]), c(s["elcex_be"]), f(s["elcam_be"]),
            f(s["elcbfc_be"]), c(s["agnbfcc"]), c(s["nucex"]), f(s["nucam"]), f(s["nucbfc"])]
    args = [C.c_int64(s[k]) for k in ("nebf", "npebf", "nebfBE", "npebfBE", "npbf")] + [_p(a) for a in keep]
    return args, keep


def _mdims(s):
    return s["nebf"], s["nebfBE"], s["npbf"]


def rxchfmult_gam2(s):
    na, nb, npb = _mdims(s)
    gi = np.zeros(npb * npb * na * na * nb * nb)
    gx, gs = np.zeros(npb * npb * nb ** 4), np.zeros(npb * npb * nb ** 4)
    ba, k1 = _mbas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam2(*ba, *sa, _p(gi), _p(gx), _p(gs))
    return gi, gx, gs


def rxchfmult_gam3_int(s):
    na, nb, npb = _mdims(s)
    g = np.zeros(npb * npb * na * na * nb ** 4)
    ba, k1 = _mbas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam3_int(*ba, *sa, _p(g))
    return g


def rxchfmult_gam4_int(na, nb, npb, g2s, gee):
    g = np.zeros(npb * npb * na * na * nb ** 6)
    lib().orc_rxchfmult_gam4_int(C.c_int64(na), C.c_int64(nb), C.c_int64(npb), _p(np.ascontiguousarray(g2s)),
                                 _p(np.ascontiguousarray(gee)), _p(g))
    return g


def rxchfmult_omg(order, which, na, nb, npb, DAE, DBE, DP, G, F):
    """RXCHFmult_{FAE,FP,FBE}_OMG{order}; which 0/1/2; F accumulates."""
    e = np.zeros(1)
    lib().orc_rxchfmult_omg(C.c_int64(order), C.c_int64(which), C.c_int64(na), C.c_int64(nb), C.c_int64(npb),
                            _f(DAE), _f(DBE), _f(DP), _p(np.ascontiguousarray(G)), _p(F), _p(e))
    return float(e[0])


def _ref_mbas(s):
    c, f = np.ascontiguousarray, np.asfortranarray
    return [c(s["kpestr"]), c(s["kpeend"]), c(s["ampeb2c"]), c(s["agebfcc"]), c(s["elcex"]), f(s["elcam"]),
            f(s["elcbfc"]), c(s["kpestr_be"]), c(s["kpeend_be"]), c(s["ampeb2c_be"]), c(s["agebfcc_be"]),
            c(s["elcex_be"]), f(s["elcam_be"]), f(s["elcbfc_be"]), c(s["agnbfcc"]), c(s["nucex"]), f(s["nucam"]),
            f(s["nucbfc"])]


def ref_rxchfmult_gam2(s):
    """RXCHFmult_GAM2 (source/rxchfmult/rxchfmult_gam2.f:2)."""
    na, nb, npb = _mdims(s)
    nint, nx = npb * npb * na * na * nb * nb, npb * npb * nb ** 4
    gi, gx, gs = np.zeros(nint), np.zeros(nx), np.zeros(nx)
    keep = _ref_mbas(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfmult_gam2_(_i8(1), _i8(na), _i8(s["npebf"]), _i8(nb), _i8(s["npebfBE"]), _i8(npb), _i8(nint),
                          _i8(nx), _i8(s["nat"]), _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm),
                          *[_p(a) for a in keep], _p(gi), _p(gx), _p(gs))
    return gi, gx, gs


def ref_rxchfmult_gam3_int(s):
    """RXCHFmult_GAM3_INT (source/rxchfmult/rxchfmult_gam3-int.f:2)."""
    na, nb, npb = _mdims(s)
    n3 = npb * npb * na * na * nb ** 4
    g = np.zeros(n3)
    keep = _ref_mbas(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfmult_gam3_int_(_i8(1), _i8(na), _i8(s["npebf"]), _i8(nb), _i8(s["npebfBE"]), _i8(npb), _i8(n3),
                              _i8(s["nat"]), _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm),
                              *[_p(a) for a in keep], _p(g))
    return g


def ref_rxchfmult_gam4_int(na, nb, npb, g2s, gee, workdir):
    """RXCHFmult_GAM4_INT (source/rxchfmult/rxchfmult_gam4-int.f:2): reads
    GAM_ee.ufm from workdir."""
    np.ascontiguousarray(gee, dtype=np.float64).tofile(os.path.join(workdir, "GAM_ee.ufm"))
    g4 = np.zeros(npb * npb * na * na * nb ** 6)
    _in_dir(workdir, lambda: ref().rxchfmult_gam4_int_(_i8(1), _i8(na), _i8(nb), _i8(npb), _i8(na ** 4),
                                                       _i8(npb * npb * nb ** 4), _i8(g4.size),
                                                       _p(np.ascontiguousarray(g2s)), _p(g4)))
    return g4


def ref_rxchfmult_omg(order, which, na, nb, npb, DAE, DBE, DP, G, F):
    e = C.c_double()
    fn = getattr(ref(), f"rxchfmult_{('fae', 'fp', 'fbe')[which]}_omg{order}_")
    fn(_i8(1), _i8(na), _i8(nb), _i8(npb), _i8(np.asarray(G).size), _f(DAE), _f(DBE), _f(DP),
       _p(np.ascontiguousarray(G)), _p(F), C.byref(e))
    return e.value


def rxchfmult_gam1(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne * ne * npb * npb)
    gs = np.zeros_like(g)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam1(*ba, *sa, _p(g), _p(gs))
    return g, gs


def rxchfmult_gam3_xchf(s):
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne ** 6 * npb * npb)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam3_xchf(*ba, *sa, _p(g))
    return g


def ref_rxchfmult_gam1(s, workdir):
    """RXCHFmult_GAM1_MD (source/rxchfmult/rxchfmult_gam1.f:2): writes
    GAM_1.ufm and GAM_1s.ufm into workdir."""
    ne, npb = s["nebf"], s["npbf"]
    keep = [np.ascontiguousarray(s[k]) for k in ("ampeb2c", "agebfcc", "agnbfcc", "elcex", "nucex")]
    keep += [np.asfortranarray(s[k]) for k in ("elcam", "nucam", "elcbfc", "nucbfc")]
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    _in_dir(workdir, lambda: ref().rxchfmult_gam1_md_(_i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ne * ne * npb * npb),
                                                      _i8(s["npebf"] ** 2 * npb * npb), _i8(s["nat"]),
                                                      _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b),
                                                      _p(gm), *[_p(a) for a in keep]))
    g = np.fromfile(os.path.join(workdir, "GAM_1.ufm"), dtype=np.float64)
    gs = np.fromfile(os.path.join(workdir, "GAM_1s.ufm"), dtype=np.float64)
    return g, gs


def ref_rxchfmult_gam3_xchf(s):
    """RXCHFmult_GAM3_XCHF (source/rxchfmult/rxchfmult_gam3-xchf.f:2)."""
    ne, npb = s["nebf"], s["npbf"]
    ng3 = ne ** 6 * npb * npb
    g = np.zeros(ng3)
    c, f = np.ascontiguousarray, np.asfortranarray
    keep = [c(s["kpestr"]), c(s["kpeend"]), c(s["ampeb2c"]), c(s["agebfcc"]), c(s["elcex"]), f(s["elcam"]),
            f(s["elcbfc"]), c(s["agnbfcc"]), c(s["nucex"]), f(s["nucam"]), f(s["nucbfc"])]
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfmult_gam3_xchf_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng3), _i8(s["nat"]), _i8(s["ngtg"]),
                               _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep], _p(g))
    return g


def ref_rxchfmult_gam4_xchf(na, nb, npb, g2s, gee, workdir):
    """RXCHFmult_GAM4_XCHF (source/rxchfmult/rxchfmult_gam4-xchf.f:2): reads
    GAM_ee.ufm (regular basis) from workdir."""
    np.ascontiguousarray(gee, dtype=np.float64).tofile(os.path.join(workdir, "GAM_ee.ufm"))
    g4 = np.zeros(npb * npb * nb ** 8)
    _in_dir(workdir, lambda: ref().rxchfmult_gam4_xchf_(_i8(1), _i8(na), _i8(nb), _i8(npb), _i8(na ** 4),
                                                        _i8(npb * npb * nb ** 4), _i8(g4.size),
                                                        _p(np.ascontiguousarray(g2s)), _p(g4)))
    return g4


def rxchfmult_gam2_ic1ex(s):
    """GM2_1ICR, GM2_2ICR, GM2_3ICR, GM2sICR of RXCHFmult_GAM2_IC1ex (one basis)."""
    ne, npb = s["nebf"], s["npbf"]
    out = [np.zeros(ne ** 4 * npb * npb) for _ in range(4)]
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam2_ic1ex(*ba, *sa, *[_p(a) for a in out])
    return out


def rxchfmult_gam3_ic1ex(s):
    ne, npb = s["nebf"], s["npbf"]
    out = [np.zeros(ne ** 6 * npb * npb) for _ in range(3)]
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchfmult_gam3_ic1ex(*ba, *sa, *[_p(a) for a in out])
    return out


def rxchfmult_omgex(order, which, ne, npb, DAE, DBE, DP, G, GX1, GX2, F):
    e = np.zeros(1)
    if order == 2:
        lib().orc_rxchfmult_omg2ex(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                                   _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(GX1)), _p(F), _p(e))
    else:
        lib().orc_rxchfmult_omg3ex(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                                   _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(GX1)),
                                   _p(np.ascontiguousarray(GX2)), _p(F), _p(e))
    return float(e[0])


def ref_rxchfmult_gam2_ic1ex(s):
    """RXCHFmult_GAM2_IC1ex (source/rxchfmult/rxchfmult_gam2ex.f:2)."""
    ne, npb = s["nebf"], s["npbf"]
    ng2 = ne ** 4 * npb * npb
    out = [np.zeros(ng2) for _ in range(4)]
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfmult_gam2_ic1ex_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng2), _i8(0), _i8(s["nat"]),
                                _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm),
                                *[_p(a) for a in keep], *[_p(a) for a in out])
    return out


def ref_rxchfmult_gam3_ic1ex(s):
    """RXCHFmult_GAM3_IC1ex (source/rxchfmult/rxchfmult_gam3ex.f:2)."""
    ne, npb = s["nebf"], s["npbf"]
    ng3 = ne ** 6 * npb * npb
    out = [np.zeros(ng3) for _ in range(3)]
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchfmult_gam3_ic1ex_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng3), _i8(0), _i8(s["nat"]),
                                _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm),
                                *[_p(a) for a in keep], *[_p(a) for a in out])
    return out


def ref_rxchfmult_omgex(order, which, ne, npb, DAE, DBE, DP, G, GX1, GX2, F):
    e = C.c_double()
    fn = getattr(ref(), f"rxchfmult_{('fae', 'fp', 'fbe')[which]}_omg{order}ex_")
    gs = [_p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(GX1))]
    if order == 3:
        gs.append(_p(np.ascontiguousarray(GX2)))
    fn(_i8(1), _i8(ne), _i8(npb), _i8(np.asarray(G).size), _f(DAE), _f(DBE), _f(DP), *gs, _p(F), C.byref(e))
    return e.value


# ----------------------------------------------------------------------
# NEO-XCHF RHF SCF side (restatement in neorx_oracle.c)
# ----------------------------------------------------------------------
def enuc(s):
    """class_nuc_rep (source/xchf1_independent/class_nuc_rep.f:1)."""
    f = lib().orc_enuc
    f.restype = C.c_double
    return f(C.c_int64(s["nat"]), _p(np.ascontiguousarray(s["cat"])), _p(np.ascontiguousarray(s["zan"])))


def one_body(s):
    """calc_GAM_epcore + elec_ovlap / nuc_ovlap: (GAM_ecore, S_e, GAM_pcore,
    S_p) in pack_2D order."""
    ne, npb = s["nebf"], s["npbf"]
    He, Se, Hp, Sp = np.zeros(ne * ne), np.zeros(ne * ne), np.zeros(npb * npb), np.zeros(npb * npb)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_one_body(*ba, *sa, _p(He), _p(Se), _p(Hp), _p(Sp))
    return He, Se, Hp, Sp


def gam_ep(s):
    """calc_GAM_ep (source/xchf1_independent/gam_ep.f:1): pack_4D order."""
    ne, npb = s["nebf"], s["npbf"]
    g = np.zeros(ne * ne * npb * npb)
    ba, k1 = _bas(s)
    lib().orc_gam_ep(*ba, _p(g))
    return g


def xchf_ints(s, gam4=True):
    """Every integral of the XCHF energy for system s (oracle restatement):
    dict with enuc, He, Se, Hp, Sp, Gep, Gee, G1, G1s, G2, G2s, G3, G4
    (G4: the in-core 5-particle tensor, or None)."""
    He, Se, Hp, Sp = one_body(s)
    ints = dict(enuc=enuc(s), He=He, Se=Se, Hp=Hp, Sp=Sp, Gep=gam_ep(s), Gee=gam_ee(s))
    ints["G1"], ints["G1s"] = gam1(s)
    ints["G2"], ints["G2s"] = gam2_ic1(s)
    ints["G3"] = gam3_ic1(s) if s.get("nelec", 3) > 2 else np.zeros(1)
    ints["G4"] = gam4_icr(s["nebf"], s["npbf"], ints["G2s"], ints["Gee"]) if gam4 else None
    return ints


def _ints_args(ints):
    keep = [np.ascontiguousarray(ints[k], dtype=np.float64) for k in ("He", "Hp", "Gep", "Gee", "G1", "G1s", "G2",
                                                                      "G2s", "G3")]
    g4 = None if ints.get("G4") is None else np.ascontiguousarray(ints["G4"])
    return [_p(a) for a in keep] + [None if g4 is None else _p(g4)], keep + [g4]


def xchf_fock(nebf, npbf, nelec, ints, DE, DP):
    """xchf1_fock (source/xchf1_independent/xchf1_fock.f:2): (FE, FP, comps)."""
    FE, FP, comps = np.zeros((nebf, nebf), order="F"), np.zeros((npbf, npbf), order="F"), np.zeros(14)
    a, keep = _ints_args(ints)
    de, dp = np.asfortranarray(DE), np.asfortranarray(DP)
    lib().orc_xchf_fock(C.c_int64(nebf), C.c_int64(npbf), C.c_int64(nelec), C.c_double(ints["enuc"]), _p(de), _p(dp),
                        *a, _p(FE), _p(FP), _p(comps))
    return FE, FP, comps


def roothan(S, F, kstart, klast, prefac):
    """ROOTHAN (source/xchf1_independent/xcscf.f:766): (C, eigenvalues, D)."""
    n = S.shape[0]
    Cm, ev, D = np.zeros((n, n), order="F"), np.zeros(n), np.zeros((n, n), order="F")
    s_, f_ = np.asfortranarray(S), np.asfortranarray(F)
    lib().orc_roothan(C.c_int64(n), _p(s_), _p(f_), C.c_int64(kstart), C.c_int64(klast), C.c_double(prefac), _p(Cm),
                      _p(ev), _p(D))
    return Cm, ev, D


def xcscf(s, ints, ce_guess=None, maxit=100, tole=1e-6, tolp=1e-4):
    """The SCF loop of xcscf (source/xchf1_independent/xcscf.f:2) with ROOTHAN
    steps only: dict(energy, niter, comps, comps_init, trace)."""
    ne, npb = s["nebf"], s["npbf"]
    a, keep = _ints_args(ints)
    ci, cf, tr = np.zeros(14), np.zeros(14), np.zeros((maxit, 4))
    ce = None if ce_guess is None else np.asfortranarray(ce_guess, dtype=np.float64)
    se, sp = (np.ascontiguousarray(ints[k]) for k in ("Se", "Sp"))
    # Se/Sp are symmetric: pack_2D order equals column-major
    f = lib().orc_xcscf
    f.restype = C.c_int64
    n = f(C.c_int64(ne), C.c_int64(npb), C.c_int64(s["nelec"]), C.c_int64(s.get("nucst", 1)),
          C.c_double(ints["enuc"]), _p(se), _p(sp), *a, None if ce is None else _p(ce), C.c_int64(maxit),
          C.c_double(tole), C.c_double(tolp), _p(ci), _p(cf), _p(tr))
    return dict(energy=cf[0], niter=int(n), comps=cf, comps_init=ci, trace=tr[:abs(int(n))])


# ----------------------------------------------------------------------
# The reference's own NEO-XCHF SCF (xcscf, source/xchf1_independent/xcscf.f:2)
# run from oracle/_ref/libref.so in a child process: the integral routines
# of xcneo_HYBRID.f:663-1002 write their .ufm files into a work directory,
# the in-core GM2/GM3/GM4 are passed as arrays, and xcscf's printed log is
# parsed.  Used to make golden vectors (tests/golden/make_xchf_scf.py).
# ----------------------------------------------------------------------
_XCSCF_CHILD = r"""
import os, sys, json
import numpy as np
sys.path.insert(0, {odir!r})
import oracle
s = {{k: v for k, v in np.load("system.npz").items()}}
for k in ("nebf", "npebf", "npbf", "nat", "ngtg", "nelec", "nucst"):
    s[k] = int(s[k])
s["pmass"] = float(s["pmass"])
fd = os.open("xcscf.log", os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
os.dup2(fd, 1)
oracle._ref_xcscf_here(s, {read_ce}, {lsoscf})
sys.stdout.flush()
"""


def _ref_xcscf_here(s, read_ce, lsoscf):
    R = ref()
    ne, npe, npb, nat, ngtg, nelec = (s[k] for k in ("nebf", "npebf", "npbf", "nat", "ngtg", "nelec"))
    ne2, np2 = ne * ne, npb * npb
    ng1, ng2, ng3, ng4 = ne2 * np2, ne2 * ne2 * np2, ne2 ** 3 * np2, ne2 ** 4 * np2
    cat, zan, b, g = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    kpestr, kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc = _ref_basis_args(s)
    wd = os.getcwd()
    R.class_nuc_rep_(_i8(nat), _p(zan), _p(cat))
    R.elec_ovlap_(_i8(npe), _i8(ne), _i8(ne2), _p(ampeb2c), _p(agebfcc), _p(elcex), _p(elcam), _p(elcbfc))
    R.nuc_ovlap_(_i8(npb), _i8(np2), _p(agnbfcc), _p(nucex), _p(nucam), _p(nucbfc))
    R.calc_gam_epcore_(_i8(ne), _i8(npe), _i8(npb), _i8(ne2), _i8(np2), _i8(nat), _d(s["pmass"]), _p(zan), _p(cat),
                       _p(ampeb2c), _p(agebfcc), _p(agnbfcc), _p(elcex), _p(nucex), _p(elcam), _p(nucam), _p(elcbfc),
                       _p(nucbfc))
    R.calc_gam_ep_(_i8(ne), _i8(npe), _i8(npb), _i8(ng1), _p(ampeb2c), _p(agebfcc), _p(agnbfcc), _p(elcex),
                   _p(nucex), _p(elcam), _p(nucam), _p(elcbfc), _p(nucbfc))
    ref_gam_ee(s, wd)
    ref_gam1(s, wd)
    g2, g2s = ref_gam2_ic1(s)
    g3 = ref_gam3_ic1(s) if nelec > 2 else np.zeros(1)
    g4 = ref_gam4_icr(ne, npb, g2s, np.fromfile("GAM_ee.ufm", dtype=np.float64), wd) if nelec > 2 else np.zeros(1)
    nocc = nelec // 2
    npr = (ne - nocc) * nocc
    T, F = _i8(1), _i8(0)
    R.xcscf_(_i8(nelec), _i8(npr), _i8(ne * (ne + 1) // 2), _i8(s.get("nucst", 1)), _i8(npe), _i8(ne), _i8(ne2),
             _i8(npb), _i8(np2), _i8(ne2 * ne2), _i8(ngtg), _i8(ng1), _i8(ng2), _i8(ng3), _i8(ng4), _i8(1), _i8(1),
             _i8(1), T if read_ce else F, F, F, T, F, F, F, T if lsoscf else F, F,
             _i8(npe ** 4 * np2), _i8(npe ** 6 * np2), _i8(nat), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(g),
             _p(kpestr), _p(kpeend), _p(ampeb2c), _p(agebfcc), _p(agnbfcc), _p(elcex), _p(nucex), _p(elcam),
             _p(nucam), _p(elcbfc), _p(nucbfc), T, _i8(ng2), _p(g2), _p(g2s), T, _i8(g3.size), _p(g3), T,
             _i8(g4.size), _p(g4))


def ref_one_body(s, workdir):
    """The reference's class_nuc_rep, elec_ovlap / nuc_ovlap, calc_GAM_epcore
    and calc_GAM_ep, which write ENUCRP.dat, eovlap.ufm, novlap.ufm,
    GAM_ecr.ufm, GAM_pcr.ufm and GAM_ep.ufm into workdir: returns (E_nuc,
    GAM_ecore, S_e, GAM_pcore, S_p, GAM_ep) in the files' (pack) order."""
    R = ref()
    ne, npe, npb, nat = s["nebf"], s["npebf"], s["npbf"], s["nat"]
    ne2, np2 = ne * ne, npb * npb
    cat, zan = np.ascontiguousarray(s["cat"]), np.ascontiguousarray(s["zan"])
    kpestr, kpeend, ampeb2c, agebfcc, agnbfcc, elcex, nucex, elcam, nucam, elcbfc, nucbfc = _ref_basis_args(s)

    def run():
        R.class_nuc_rep_(_i8(nat), _p(zan), _p(cat))
        R.elec_ovlap_(_i8(npe), _i8(ne), _i8(ne2), _p(ampeb2c), _p(agebfcc), _p(elcex), _p(elcam), _p(elcbfc))
        R.nuc_ovlap_(_i8(npb), _i8(np2), _p(agnbfcc), _p(nucex), _p(nucam), _p(nucbfc))
        R.calc_gam_epcore_(_i8(ne), _i8(npe), _i8(npb), _i8(ne2), _i8(np2), _i8(nat), _d(s["pmass"]), _p(zan),
                           _p(cat), _p(ampeb2c), _p(agebfcc), _p(agnbfcc), _p(elcex), _p(nucex), _p(elcam),
                           _p(nucam), _p(elcbfc), _p(nucbfc))
        R.calc_gam_ep_(_i8(ne), _i8(npe), _i8(npb), _i8(ne2 * np2), _p(ampeb2c), _p(agebfcc), _p(agnbfcc),
                       _p(elcex), _p(nucex), _p(elcam), _p(nucam), _p(elcbfc), _p(nucbfc))

    _in_dir(workdir, run)
    rd = lambda f: np.fromfile(os.path.join(workdir, f), dtype=np.float64)  # noqa: E731
    e = float(open(os.path.join(workdir, "ENUCRP.dat")).read().split()[0].replace("D", "E"))
    return e, rd("GAM_ecr.ufm"), rd("eovlap.ufm"), rd("GAM_pcr.ufm"), rd("novlap.ufm"), rd("GAM_ep.ufm")


def ref_roothan(S, F, n_particle, i_particle, nucst=1):
    """ROOTHAN (source/xchf1_independent/xcscf.f:766) of the reference:
    (D, C, eigenvalues)."""
    n = S.shape[0]
    D, Cm, ev = np.zeros((n, n), order="F"), np.zeros((n, n), order="F"), np.zeros(n)
    s_, f_ = np.asfortranarray(S, dtype=np.float64), np.asfortranarray(F, dtype=np.float64)
    ref().roothan_(_p(D), _p(Cm), _p(ev), _p(s_), _p(f_), _i8(n), _i8(n_particle), _i8(i_particle), _i8(nucst))
    return D, Cm, ev


def ref_xchf_fock(s, ints, DE, DP, workdir):
    """xchf1_fock (source/xchf1_independent/xchf1_fock.f:2) of the reference on
    the in-core path (LG2IC1, LG3IC1, LG4IC) with the given integrals; GAM_1 and
    GAM_1s go to workdir as the .ufm files E_from_GAM_1 / S_from_GAM_1s read.
    Returns (FE, FP, comps) in the order of the C ABI's comps array."""
    R = ref()
    ne, npe, npb, nat, ngtg, nelec = (s[k] for k in ("nebf", "npebf", "npbf", "nat", "ngtg", "nelec"))
    ne2, np2 = ne * ne, npb * npb
    np.ascontiguousarray(ints["G1"]).tofile(os.path.join(workdir, "GAM_1.ufm"))
    np.ascontiguousarray(ints["G1s"]).tofile(os.path.join(workdir, "GAM_1s.ufm"))
    cat, zan, b, g = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    basis = _ref_basis_args(s)
    arr = {k: np.ascontiguousarray(ints[k], dtype=np.float64) for k in ("He", "Hp", "Gep", "Gee", "G2", "G2s", "G3",
                                                                        "G4")}
    FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
    XFE, XFP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
    de, dp = np.asfortranarray(DE, dtype=np.float64), np.asfortranarray(DP, dtype=np.float64)
    out = [C.c_double() for _ in range(13)]
    out[1].value = ints["enuc"]
    T, F = _i8(1), _i8(0)
    ng1, ng2, ng3, ng4 = ne2 * np2, ne2 * ne2 * np2, arr["G3"].size, arr["G4"].size

    def run():
        R.xchf1_fock_(F, T, F, T, F, T, F, T, F, _i8(1), _i8(1), _i8(1), _i8(ng4), _i8(ng3), _i8(ng2), _i8(npe),
                      _i8(ne), _i8(ne2), _i8(npb), _i8(np2), _i8(nelec), _i8(ne2 * ne2), _i8(ng1), _i8(ng2),
                      _i8(ng3), _i8(ng4), _p(de), _p(dp), _p(arr["He"]), _p(arr["Hp"]), _p(arr["Gep"]),
                      _p(arr["Gee"]), _p(arr["G4"]), _p(arr["G3"]), _p(arr["G2"]), _p(arr["G2s"]),
                      _i8(npe ** 4 * np2), _i8(ngtg), _i8(npe ** 6 * np2), _i8(nat), _d(s["pmass"]), _p(cat),
                      _p(zan), _p(b), _p(g), *[_p(a) for a in basis], _p(FE), _p(FP), _p(XFE), _p(XFP),
                      *[C.byref(o) for o in out])

    _in_dir(workdir, run)
    v = [o.value for o in out]  # E_total E_nuc E_ecore E_pcore E_ep E_ee E_gam1..4 S_total S_gam1 S_gam2
    psi = v[0] - v[1]
    comps = np.array(v[:10] + [v[10], v[11], v[12], psi / v[10]])
    return FE, FP, comps


_COMP_KEYS = ("E_NUC", "E_ECORE", "E_PCORE", "E_EP", "E_EE", "E_GAM1", "E_GAM2", "E_GAM3", "E_GAM4", "S_TOTAL",
              "E_TOTAL")


def parse_xcscf_log(text):
    """Initial-guess and final energy components, the iteration table and the
    final energy of an xcscf log (the formats 9000-9300 of xcscf.f)."""
    import re
    blocks = [dict(re.findall(r"^\s*(E_[A-Z0-9]+|S_TOTAL)=\s*(\S+)", b, re.M))
              for b in text.split("NEOXCHF ENERGETIC COMPONENTS")[1:]]
    comps = [{k: float(v) for k, v in b.items()} for b in blocks]
    its = [tuple(float(x) for x in m) for m in
           re.findall(r"^\s+(\d+)\s+(-?\d+\.\d+)\s+(-?\d+\.\d+)\s+(\d+\.\d+)\s+(\d+\.\d+)\s+(\d+\.\d+)\s*$", text, re.M)]
    m = re.search(r"FINAL NEOXCHF ENERGY IS\s+(\S+)\s+AFTER\s+(\d+)", text)
    return dict(initial=comps[0] if comps else None, final=comps[1] if len(comps) > 1 else None,
                iterations=its, energy=float(m.group(1)) if m else None, niter=int(m.group(2)) if m else None)


def ref_xcscf(s, workdir, ce_guess=None, lsoscf=True, timeout=3600):
    """Run the reference's integrals + xcscf for system dict s (with nelec,
    nucst) in workdir (a child process; OMP threads as the environment sets)
    and return parse_xcscf_log of its output.  ce_guess: orbitals written as
    guessCE.inp (read_CE = .true.), else the core guess."""
    import subprocess
    import sys
    os.makedirs(workdir, exist_ok=True)
    np.savez(os.path.join(workdir, "system.npz"), **{k: np.asarray(v) for k, v in s.items() if k != "name"})
    if ce_guess is not None:
        n = ce_guess.shape[0]
        lines = []
        for j in range(n):
            for ic, i0 in enumerate(range(0, n, 5)):
                vals = "".join(f"{v:15.8E}" for v in ce_guess[i0:i0 + 5, j])
                lines.append(f"{(j + 1) % 100:2d}{(ic + 1) % 1000:3d}{vals}")
        open(os.path.join(workdir, "guessCE.inp"), "w").write("\n".join(lines) + "\n")
    code = _XCSCF_CHILD.format(odir=os.path.dirname(os.path.abspath(__file__)), read_ce=ce_guess is not None,
                               lsoscf=lsoscf)
    subprocess.run([sys.executable, "-c", code], cwd=workdir, check=True, timeout=timeout)
    return parse_xcscf_log(open(os.path.join(workdir, "xcscf.log")).read())


# ---------------------------------------------------------------------------
# RXCHF, one special electron (oracle and reference)
# ---------------------------------------------------------------------------
def rxchf_gam2_ic1(s):
    """RXCHF_GAM2_IC1 restated: (GM2_1ICR, GM2_2ICR, GM2sICR)."""
    ne, npb = s["nebf"], s["npbf"]
    g1 = np.zeros(ne ** 4 * npb * npb)
    g2, gs = np.zeros_like(g1), np.zeros_like(g1)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchf_gam2_ic1(*ba, *sa, _p(g1), _p(g2), _p(gs))
    return g1, g2, gs


def rxchf_omg2(which, ne, npb, DAE, DBE, DP, G1, G2, Gs, F, SF):
    """which: 0 RXCHF_FAE_OMG2, 1 RXCHF_FP_OMG2, 2 RXCHF_FBE_OMG2; F and SF
    accumulate in place; returns (E, S)."""
    e, sv = np.zeros(1), np.zeros(1)
    lib().orc_rxchf_omg2(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                         _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G2)), _p(np.ascontiguousarray(Gs)),
                         _p(F), _p(SF), _p(e), _p(sv))
    return float(e[0]), float(sv[0])


def ref_rxchf_gam2_ic1(s):
    """RXCHF_GAM2_IC1 (source/rxchf/rxchf_gam2_IC1_OMP.f:2) of the reference."""
    ne, npb = s["nebf"], s["npbf"]
    ng2 = ne ** 4 * npb * npb
    g1, g2, gs = np.zeros(ng2), np.zeros(ng2), np.zeros(ng2)
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchf_gam2_ic1_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng2), _i8(0), _i8(s["nat"]),
                          _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep],
                          _p(g1), _p(g2), _p(gs))
    return g1, g2, gs


def ref_rxchf_omg2(which, ne, npb, DAE, DBE, DP, G1, G2, Gs, F, SF):
    """RXCHF_FAE_OMG2 / _FP_OMG2 / _FBE_OMG2 (source/rxchf/rxchf_omg2.f:2,
    :231, :449) of the reference; returns (E, S)."""
    e, sv = C.c_double(), C.c_double()
    ng2 = ne ** 4 * npb * npb
    R = ref()
    fn = (R.rxchf_fae_omg2_, R.rxchf_fp_omg2_, R.rxchf_fbe_omg2_)[which]
    fn(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _f(DAE), _f(DBE), _f(DP), _p(np.ascontiguousarray(G1)),
       _p(np.ascontiguousarray(G2)), _p(np.ascontiguousarray(Gs)), _p(F), _p(SF), C.byref(e), C.byref(sv))
    return e.value, sv.value


def rxchf_gam3_ic1(s):
    """RXCHF_GAM3_IC1 restated: (GM3_1ICR, GM3_2ICR)."""
    ne, npb = s["nebf"], s["npbf"]
    g1 = np.zeros(ne ** 6 * npb * npb)
    g2 = np.zeros_like(g1)
    ba, k1 = _bas(s)
    sa, k2 = _sysargs(s)
    lib().orc_rxchf_gam3_ic1(*ba, *sa, _p(g1), _p(g2))
    return g1, g2


def rxchf_omg3(which, ne, npb, DAE, DBE, DP, G1, G2, F):
    """which: 0 RXCHF_FAE_OMG3, 1 RXCHF_FP_OMG3, 2 RXCHF_FBE_OMG3; F
    accumulates in place; returns E."""
    e = np.zeros(1)
    lib().orc_rxchf_omg3(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAE), _f(DBE), _f(DP),
                         _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G2)), _p(F), _p(e))
    return float(e[0])


def rxchf_omg1(which, ne, npb, DBE, DP, G1, G1s, F, SF):
    """which: 1 RXCHF_FP_OMG1, 2 RXCHF_FBE_OMG1 over the GAM_1 / GAM_1s
    arrays; F accumulates, SF is overwritten; returns (E, S)."""
    e, sv = np.zeros(1), np.zeros(1)
    lib().orc_rxchf_omg1(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DBE), _f(DP),
                         _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G1s)), _p(F), _p(SF), _p(e), _p(sv))
    return float(e[0]), float(sv[0])


def ref_rxchf_gam3_ic1(s):
    """RXCHF_GAM3_IC1 (source/rxchf/rxchf_gam3_IC1_OMP.f:2) of the reference."""
    ne, npb = s["nebf"], s["npbf"]
    ng3 = ne ** 6 * npb * npb
    g1, g2 = np.zeros(ng3), np.zeros(ng3)
    keep = _ref_basis_args(s)
    cat, zan, b, gm = (np.ascontiguousarray(s[k]) for k in ("cat", "zan", "bcoef", "gamma"))
    ref().rxchf_gam3_ic1_(_i8(1), _i8(ne), _i8(s["npebf"]), _i8(npb), _i8(ng3), _i8(0), _i8(s["nat"]),
                          _i8(s["ngtg"]), _d(s["pmass"]), _p(cat), _p(zan), _p(b), _p(gm), *[_p(a) for a in keep],
                          _p(g1), _p(g2))
    return g1, g2


def ref_rxchf_omg3(which, ne, npb, DAE, DBE, DP, G1, G2, F):
    """RXCHF_FAE_OMG3 / _FP_OMG3 / _FBE_OMG3 (source/rxchf/rxchf_omg3.f:2,
    :263, :486) of the reference; returns E."""
    e = C.c_double()
    ng3 = ne ** 6 * npb * npb
    R = ref()
    fn = (R.rxchf_fae_omg3_, R.rxchf_fp_omg3_, R.rxchf_fbe_omg3_)[which]
    fn(_i8(1), _i8(ne), _i8(npb), _i8(ng3), _f(DAE), _f(DBE), _f(DP), _p(np.ascontiguousarray(G1)),
       _p(np.ascontiguousarray(G2)), _p(F), C.byref(e))
    return e.value


def ref_rxchf_omg1(which, ne, npb, DBE, DP, G1, G1s, F, SF, workdir):
    """RXCHF_FP_OMG1 / RXCHF_FBE_OMG1 (source/rxchf/rxchf_omg1.f:2, :84) of
    the reference: they read GAM_1.ufm and GAM_1s.ufm from the working
    directory (8-byte records, the flang build), written here into workdir
    first; returns (E, S)."""
    np.ascontiguousarray(G1, dtype=np.float64).tofile(os.path.join(workdir, "GAM_1.ufm"))
    np.ascontiguousarray(G1s, dtype=np.float64).tofile(os.path.join(workdir, "GAM_1s.ufm"))
    e, sv = C.c_double(), C.c_double()
    R = ref()
    fn = R.rxchf_fp_omg1_ if which == 1 else R.rxchf_fbe_omg1_
    cwd = os.getcwd()
    os.chdir(workdir)
    try:
        fn(_i8(ne), _i8(npb), _i8(len(G1)), _f(DBE), _f(DP), _p(F), _p(SF), C.byref(e), C.byref(sv))
    finally:
        os.chdir(cwd)
    return e.value, sv.value


def rxcuhf_omg2(which, ne, npb, DAalp, DAtot, DBE, DP, G1, G2, Gs, F, SF, F2=None):
    """which: 0 RXCUHF_FE_OMG2 (F = FAalpE, F2 = FAbetE, SF = SAalpE),
    1 RXCUHF_FP_OMG2, 2 RXCUHF_FBE_OMG2; F, F2, SF accumulate; returns (E, S)."""
    e, sv = np.zeros(1), np.zeros(1)
    lib().orc_rxcuhf_omg2(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAalp), _f(DAtot), _f(DBE), _f(DP),
                          _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G2)), _p(np.ascontiguousarray(Gs)),
                          _p(F), None if F2 is None else _p(F2), _p(SF), _p(e), _p(sv))
    return float(e[0]), float(sv[0])


def rxcuhf_omg3(which, ne, npb, DAalp, DAbet, DAtot, DBE, DP, G1, G2, F, F2=None):
    """which: 0 RXCUHF_FE_OMG3 (F = FAalpE, F2 = FAbetE), 1 RXCUHF_FP_OMG3,
    2 RXCUHF_FBE_OMG3; F, F2 accumulate; returns E."""
    e = np.zeros(1)
    lib().orc_rxcuhf_omg3(C.c_int64(which), C.c_int64(ne), C.c_int64(npb), _f(DAalp), _f(DAbet), _f(DAtot), _f(DBE),
                          _f(DP), _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G2)), _p(F),
                          None if F2 is None else _p(F2), _p(e))
    return float(e[0])


def ref_rxcuhf_omg2(which, ne, npb, DAalp, DAbet, DAtot, DBE, DP, G1, G2, Gs, F, SF, F2=None):
    """RXCUHF_FE_OMG2 / _FP_OMG2 / _FBE_OMG2 (source/rxchf/rxcuhf_omg2.f:2,
    :253, :469) of the reference; returns (E, S)."""
    e, sv = C.c_double(), C.c_double()
    ng2 = ne ** 4 * npb * npb
    R = ref()
    g = [_p(np.ascontiguousarray(x)) for x in (G1, G2, Gs)]
    if which == 0:
        R.rxcuhf_fe_omg2_(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _f(DAalp), _f(DAbet), _f(DAtot), _f(DBE), _f(DP), *g,
                          _p(F), _p(F2), _p(SF), C.byref(e), C.byref(sv))
    else:
        fn = R.rxcuhf_fp_omg2_ if which == 1 else R.rxcuhf_fbe_omg2_
        fn(_i8(1), _i8(ne), _i8(npb), _i8(ng2), _i8(1), _i8(1), _f(DAalp), _f(DAtot), _f(DBE), _f(DP), *g, _p(F),
           _p(SF), C.byref(e), C.byref(sv))
    return e.value, sv.value


def ref_rxcuhf_omg3(which, ne, npb, DAalp, DAbet, DAtot, DBE, DP, G1, G2, F, F2=None):
    """RXCUHF_FE_OMG3 / _FP_OMG3 / _FBE_OMG3 (source/rxchf/rxcuhf_omg3.f:2,
    :270, :499) of the reference; returns E."""
    e = C.c_double()
    ng3 = ne ** 6 * npb * npb
    R = ref()
    a = [_i8(1), _i8(ne), _i8(npb), _i8(ng3), _f(DAalp), _f(DAbet), _f(DAtot), _f(DBE), _f(DP),
         _p(np.ascontiguousarray(G1)), _p(np.ascontiguousarray(G2))]
    if which == 0:
        R.rxcuhf_fe_omg3_(*a, _p(F), _p(F2), C.byref(e))
    else:
        (R.rxcuhf_fp_omg3_ if which == 1 else R.rxcuhf_fbe_omg3_)(*a, _p(F), C.byref(e))
    return e.value
