"""Pins of the oracle (CPU): closed forms, the reference's recorded spot
checks, and -- where oracle/_ref was built -- the reference itself."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from neorx import synth
from _util import check_close, random_tuples

HERE = os.path.dirname(os.path.abspath(__file__))
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference checker oracle/_ref not built")


def test_boys_closed_forms():
    # F_0(T) = sqrt(pi/T) erf(sqrt T)/2 (series branch accurate to ~1e-13)
    for T in [1e-6, 0.01, 0.5, 1.0, 3.7, 9.9, 15.0, 19.99]:
        exact = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert abs(oracle.boys(0, T) - exact) <= 2e-13 * exact
    for m in range(12):
        assert abs(oracle.boys(m, 0.0) - 1.0 / (2 * m + 1)) < 1e-15
    # T > 20: the reference's asymptotic branch, F_m = (2m-1)!!/(2T)^m sqrt(pi/T)/2 exactly
    for T in [20.0001, 25.0, 60.0]:
        for m in range(8):
            asym = 0.88622692545275801365 / math.sqrt(T)
            for l in range(1, m + 1):
                asym *= (l - 0.5) / T
            assert oracle.boys(m, T) == pytest.approx(asym, rel=1e-15)
    # downward identity F_m = (2T F_{m+1} + e^-T)/(2m+1) on the series branch
    for T in [0.3, 4.0, 12.0]:
        for m in range(6):
            lhs = oracle.boys(m, T)
            rhs = (2 * T * oracle.boys(m + 1, T) + math.exp(-T)) / (2 * m + 1)
            assert abs(lhs - rhs) <= 1e-12 * lhs


def _gint(op, blk, GA, GB, a=0, b=1, mass=1.0, order=None, pi=0.0):
    import ctypes as C
    n = len(blk["bra"])
    order = order or list(range(n))
    br = [blk["bra"][i] for i in order]
    kt = [blk["ket"][i] for i in order]
    pw = np.array([p[:3] for p in br + kt], dtype=np.int64)
    ex = np.array([p[3] for p in br + kt])
    cen = np.array([p[4:] for p in br + kt])
    return oracle.lib().orc_gint(C.c_int64(op), C.c_int64(n), oracle._p(pw), oracle._p(ex), oracle._p(cen),
                                 oracle._p(np.ascontiguousarray(GA, dtype=float)),
                                 oracle._p(np.ascontiguousarray(GB, dtype=float)), C.c_int64(a), C.c_int64(b),
                                 oracle._p(np.array(blk["C"], float)), C.c_double(mass), C.c_double(pi))


def test_spot_check_golden_values():
    G = json.load(open(os.path.join(HERE, "golden", "spot_check.json")))
    b = G["G2"]
    g1, g2 = b["gamma1"], b["gamma2"]
    GA, GB = np.array([[0, g1], [0, 0]]), np.array([[0, g2], [0, 0]])
    assert _gint(0, b, GA, GB) == pytest.approx(b["MD"]["XGGS"], rel=1e-12)
    assert -b["Z"] * _gint(1, b, GA, GB) == pytest.approx(b["MD"]["XGGVEC"], rel=1e-12)
    assert -_gint(2, b, GA, GB) == pytest.approx(b["MD"]["XGGVEP"], rel=1e-12)
    b = G["G3"]
    g1, g2 = b["gamma1"], b["gamma2"]

    def m3(A12=0, A13=0, A23=0):
        M = np.zeros((3, 3))
        M[0, 1], M[0, 2], M[1, 2] = A12, A13, A23
        return M

    md = b["MD"]
    assert _gint(0, b, m3(A13=g1), m3(A23=g2)) == pytest.approx(md["xgSg"], rel=1e-12)
    assert _gint(3, b, m3(A23=g1), m3()) == pytest.approx(md["xgTE"], rel=1e-12)
    assert _gint(3, b, m3(A13=g1), m3(A23=g2)) == pytest.approx(md["xgTEg1"], rel=1e-12)
    assert _gint(3, b, m3(A23=g1), m3(A23=g2)) == pytest.approx(md["xgTEg3"], rel=1e-12)
    assert _gint(3, b, m3(A23=g1), m3(A13=g2)) == pytest.approx(md["xgTEg2"], rel=1e-12)
    assert _gint(3, b, m3(A12=g1), m3(A13=g2), mass=1836.0, order=[2, 0, 1]) == pytest.approx(md["xgTPg"], rel=1e-12)
    assert -b["Z"] * _gint(1, b, m3(A23=g1), m3()) == pytest.approx(md["xgVEC"], rel=1e-12)
    b = G["G4"]
    g1, g2 = b["gamma1"], b["gamma2"]
    M = np.zeros((4, 4))
    M[0, 3], M[1, 3] = g1, g2
    assert -_gint(2, b, M, np.zeros((4, 4)), a=2, b=3) == pytest.approx(b["MD"]["xgVEPg"], rel=1e-12)
    M = np.zeros((4, 4))
    M[0, 3], M[2, 3] = g1, g2
    assert _gint(2, b, M, np.zeros((4, 4)), a=0, b=1) == pytest.approx(b["MD"]["xgVEEg2"], rel=1e-12)


def test_s_type_closed_forms():
    """Geminal-coupled s-type overlap = pi^{3n/2} det(M)^{-3/2} exp(-P'QP);
    the gamma -> 0 limit of the 3-particle Coulomb is the product of an ERI
    and an overlap."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        ex = rng.uniform(0.3, 3.0, 6)
        cen = rng.uniform(-1, 1, (6, 3))
        g13, g23 = rng.uniform(0.5, 4.0, 2)
        blk = {"bra": [[0, 0, 0, ex[i], *cen[i]] for i in range(3)],
               "ket": [[0, 0, 0, ex[3 + i], *cen[3 + i]] for i in range(3)], "C": [0, 0, 0]}
        GA = np.zeros((3, 3))
        GA[0, 2], GA[1, 2] = g13, g23
        val = _gint(0, blk, GA, np.zeros((3, 3)))
        p = ex[:3] + ex[3:]
        P = (ex[:3, None] * cen[:3] + ex[3:, None] * cen[3:]) / p[:, None]
        K = np.exp(-(ex[:3] * ex[3:] / p) * np.sum((cen[:3] - cen[3:]) ** 2, axis=1)).prod()
        M = np.diag(p) + np.array([[g13, 0, -g13], [0, g23, -g23], [-g13, -g23, g13 + g23]])
        Q = np.diag(p) - np.diag(p) @ np.linalg.inv(M) @ np.diag(p)
        expo = sum(P[:, d] @ Q @ P[:, d] for d in range(3))
        closed = math.pi ** 4.5 / np.linalg.det(M) ** 1.5 * math.exp(-expo) * K
        assert val == pytest.approx(closed, rel=1e-12)
        # gamma -> 0: <1/r12> factorises
        z = np.zeros((3, 3))
        v3 = _gint(2, blk, z, z, a=0, b=1)
        b12 = {"bra": blk["bra"][:2], "ket": blk["ket"][:2], "C": [0, 0, 0]}
        b3 = {"bra": blk["bra"][2:], "ket": blk["ket"][2:], "C": [0, 0, 0]}
        eri = _gint(2, b12, np.zeros((2, 2)), np.zeros((2, 2)), a=0, b=1)
        s3 = _gint(0, b3, np.zeros((1, 1)), np.zeros((1, 1)))
        assert v3 == pytest.approx(eri * s3, rel=1e-12)


@needs_ref
def test_boys_bitwise_vs_reference():
    for T in [0.0, 1e-3, 0.7, 5.5, 19.5, 20.0, 20.5, 44.0]:
        for m in range(0, 17, 3):
            assert oracle.boys(m, T) == oracle.ref_boys(m, T)


@needs_ref
@pytest.mark.parametrize("cls,n", [(1, 120), (2, 12), (3, 2)])
def test_oracle_tuples_vs_reference(cls, n):
    s = synth.small(seed=21 + cls, e_shells=(0, 1, 2) if cls < 3 else (0, 1), p_shells=(0, 1, 2) if cls < 3 else (0, 1))
    idx = random_tuples(s, cls, n, seed=cls)
    re, rs = oracle.ref_prim_list(s, cls, idx)
    oe, os_ = oracle.prim_list(s, cls, idx)
    check_close(oe, re, what=f"oracle vs reference cls{cls} E")
    if cls < 3:
        check_close(os_, rs, what=f"oracle vs reference cls{cls} S")


@needs_ref
def test_oracle_gam2_ic1_and_eg2_vs_reference():
    s = synth.small(seed=5, n_e_centres=1, e_shells=(0,), contract=(2,), p_shells=(0,), ngtg=2)
    g, gs = oracle.gam2_ic1(s)
    rg, rgs = oracle.ref_gam2_ic1(s)
    check_close(g, rg, what="GAM2_IC1 GM2ICR")
    check_close(gs, rgs, what="GAM2_IC1 GM2sICR")
    ne, npb = s["nebf"], s["npbf"]
    rng = np.random.default_rng(1)
    DE, DP = synth.density(ne, 2.0, 1), synth.density(npb, 1.0, 2)
    G = rng.standard_normal(ne ** 4 * npb * npb)
    Gs = rng.standard_normal(ne ** 4 * npb * npb)
    FE0, FP0 = np.asfortranarray(rng.standard_normal((ne, ne))), np.asfortranarray(rng.standard_normal((npb, npb)))
    FE1, FP1, FE2, FP2 = FE0.copy(order="F"), FP0.copy(order="F"), FE0.copy(order="F"), FP0.copy(order="F")
    a = oracle.eg2ic1(ne, npb, DE, DP, G, Gs, FE1, FP1)
    b = oracle.ref_eg2ic1(ne, npb, DE, DP, G, Gs, FE2, FP2)
    check_close(FE1, FE2, what="EG2 FE")
    check_close(FP1, FP2, what="EG2 FP")
    check_close(a[0], b[0], what="EG2 SE")
    check_close(a[1], b[1], what="EG2 SP")
    assert a[2] == pytest.approx(b[2], rel=1e-12) and a[3] == pytest.approx(b[3], rel=1e-12)


@needs_ref
def test_oracle_eg3_eg4_vs_reference():
    rng = np.random.default_rng(4)
    ne, npb = 3, 2
    DE, DP = synth.density(ne, 2.0, 3), synth.density(npb, 1.0, 4)
    G3 = rng.standard_normal(ne ** 6 * npb * npb)
    FE0, FP0 = np.asfortranarray(rng.standard_normal((ne, ne))), np.asfortranarray(rng.standard_normal((npb, npb)))
    F = [FE0.copy(order="F"), FP0.copy(order="F"), FE0.copy(order="F"), FP0.copy(order="F")]
    e1 = oracle.eg3ic1(ne, npb, DE, DP, G3, F[0], F[1])
    e2 = oracle.ref_eg3ic1(ne, npb, DE, DP, G3, F[2], F[3])
    check_close(F[0], F[2], what="EG3 FE")
    check_close(F[1], F[3], what="EG3 FP")
    assert e1 == pytest.approx(e2, rel=1e-12)
    ne = 2
    DE = synth.density(ne, 2.0, 5)
    G4 = rng.standard_normal(ne ** 8 * npb * npb)
    F = [FE0[:ne, :ne].copy(order="F"), FP0.copy(order="F"), FE0[:ne, :ne].copy(order="F"), FP0.copy(order="F")]
    e1 = oracle.eg4icr(ne, npb, DE, DP, G4, F[0], F[1])
    e2 = oracle.ref_eg4icr(ne, npb, DE, DP, G4, F[2], F[3])
    check_close(F[0], F[2], what="EG4 FE")
    check_close(F[1], F[3], what="EG4 FP")
    assert e1 == pytest.approx(e2, rel=1e-12)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference checker oracle/_ref not built")
def test_oracle_direct_scf_vs_reference(tmp_path):
    """The oracle's GAM2/3/4_DSCF against the reference's own routines
    (gam2_DSCF_OMP.f, gam3_DSCF_OMP.f, gam4_DSCF_OMP.f), Fock matrices starting
    non-zero so that the accumulation semantics are pinned too."""
    s = synth.small(seed=5, n_e_centres=2, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    ne, npb = s["nebf"], s["npbf"]
    DE, DP = synth.density(ne, 2.0, 1), synth.density(npb, 1.0, 2)
    rng = np.random.default_rng(0)
    F0 = [np.asfortranarray(rng.standard_normal((n, n))) for n in (ne, npb, ne, npb)]
    A = [f.copy(order="F") for f in F0]
    B = [f.copy(order="F") for f in F0]
    eo = oracle.gam2_dscf(s, DE, DP, *A)
    er = oracle.ref_gam2_dscf(s, DE, DP, *B)
    for a, b in zip(A, B):
        check_close(a, b, what="GAM2_DSCF")
    assert eo[0] == pytest.approx(er[0], rel=1e-11) and eo[1] == pytest.approx(er[1], rel=1e-11)
    A = [f.copy(order="F") for f in F0[:2]]
    B = [f.copy(order="F") for f in F0[:2]]
    assert oracle.gam3_dscf(s, DE, DP, *A) == pytest.approx(oracle.ref_gam3_dscf(s, DE, DP, *B), rel=1e-11)
    for a, b in zip(A, B):
        check_close(a, b, what="GAM3_DSCF")
    g2, g2s = oracle.gam2_ic1(s)
    gee = oracle.gam_ee(s)
    A = [f.copy(order="F") for f in F0[:2]]
    B = [f.copy(order="F") for f in F0[:2]]
    eo4 = oracle.gam4_dscf(ne, npb, g2s, gee, DE, DP, *A)
    er4 = oracle.ref_gam4_dscf(ne, npb, g2s, gee, DE, DP, *B, workdir=str(tmp_path))
    assert eo4 == pytest.approx(er4, rel=1e-11)
    for a, b in zip(A, B):
        check_close(a, b, what="GAM4_DSCF")


@needs_ref
def test_oracle_cheap_builders_vs_reference(tmp_path):
    """GAM1_OMP_MD, calc_GAM_ee, GAM3_IC1 and GAM4_ICR of the oracle against
    the reference's own routines (files they write/read live in tmp_path)."""
    s = synth.small(seed=14, n_e_centres=2, e_shells=(0, 1), contract=(1, 2), p_shells=(0,), ngtg=2)
    ne, npb = s["nebf"], s["npbf"]
    g1, g1s = oracle.gam1(s)
    r1, r1s = oracle.ref_gam1(s, str(tmp_path))
    check_close(g1, r1, what="GAM_1")
    check_close(g1s, r1s, what="GAM_1s")
    gee = oracle.gam_ee(s)
    check_close(gee, oracle.ref_gam_ee(s, str(tmp_path)), what="GAM_ee")
    t = synth.small(seed=15, n_e_centres=1, e_shells=(0, 1), contract=(1,), p_shells=(0,), ngtg=2)
    check_close(oracle.gam3_ic1(t), oracle.ref_gam3_ic1(t), what="GAM3_IC1")
    u = synth.small(seed=16, n_e_centres=2, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    g2, g2s = oracle.gam2_ic1(u)
    ue = oracle.gam_ee(u)
    check_close(oracle.gam4_icr(u["nebf"], u["npbf"], g2s, ue),
                oracle.ref_gam4_icr(u["nebf"], u["npbf"], g2s, ue, str(tmp_path)), what="GAM4_ICR")


@needs_ref
@pytest.mark.parametrize("nae,nbe", [(1, 1), (2, 1), (3, 2), (4, 4), (2, 3)])
def test_oracle_uhf_contractions_vs_reference(nae, nbe):
    """FAE_OMG2/3/4 and UFP_OMG2/3/4 of the oracle against the reference's
    own routines (source/xcuhf/xcuhf_omg{2,3,4}.f) on random tensors, for
    electron counts that switch the spin-pair flags on and off."""
    rng = np.random.default_rng(10 * nae + nbe)
    ne, npb = 3, 2
    DAE, DBE = synth.density(ne, float(nae), 1), synth.density(ne, float(nbe), 2)
    DP = synth.density(npb, 1.0, 3)
    g2, g2s = rng.standard_normal(ne ** 4 * npb * npb), rng.standard_normal(ne ** 4 * npb * npb)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    g4 = rng.standard_normal(ne ** 8 * npb * npb)
    args = (nae, nbe, ne, npb, DAE, DBE, DP)
    for fn, rfn, G, n in ((oracle.fae_omg2, oracle.ref_fae_omg2, (g2, g2s), ne),
                          (oracle.ufp_omg2, oracle.ref_ufp_omg2, (g2, g2s), npb)):
        F0 = [np.asfortranarray(rng.standard_normal((n, n))) for _ in range(2)]
        A = [f.copy(order="F") for f in F0]
        B = [f.copy(order="F") for f in F0]
        ea, sa = fn(*args, *G, *A)
        eb, sb = rfn(*args, *G, *B)
        for a, b in zip(A, B):
            check_close(a, b, what=fn.__name__)
        assert ea == pytest.approx(eb, rel=1e-12, abs=1e-14) and sa == pytest.approx(sb, rel=1e-12, abs=1e-14)
    for fn, rfn, G, n in ((oracle.fae_omg3, oracle.ref_fae_omg3, g3, ne), (oracle.ufp_omg3, oracle.ref_ufp_omg3, g3, npb),
                          (oracle.fae_omg4, oracle.ref_fae_omg4, g4, ne), (oracle.ufp_omg4, oracle.ref_ufp_omg4, g4, npb)):
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        ea = fn(*args, G, A)
        eb = rfn(*args, G, B)
        check_close(A, B, what=fn.__name__)
        assert ea == pytest.approx(eb, rel=1e-12, abs=1e-14), fn.__name__


@needs_ref
def test_oracle_omg4_icr_vs_reference(tmp_path):
    u = synth.small(seed=16, n_e_centres=2, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    g2, g2s = oracle.gam2_ic1(u)
    ue = oracle.gam_ee(u)
    check_close(oracle.omg4_icr(u["nebf"], u["npbf"], g2s, ue),
                oracle.ref_omg4_icr(u["nebf"], u["npbf"], g2s, ue, str(tmp_path)), what="OMG4_ICR")


@needs_ref
def test_oracle_rxchfne_vs_reference():
    """RXCHF-ne: the two-output e-e-p tuple, RXCHFne_GAM2_IC1 and the
    RXCHF-ne / RXCUHF-ne contractions of the oracle against the reference's
    own routines (source/rxchfne/)."""
    s = synth.small(seed=44, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    idx = random_tuples(s, 2, 16, seed=4)
    re1, re2 = oracle.ref_prim_list(s, 4, idx)
    oe1, oe2 = oracle.prim_list(s, 4, idx)
    check_close(oe1, re1, what="RXCHFne ansE1")
    check_close(oe2, re2, what="RXCHFne ansE2")
    t = synth.small(seed=45, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    g, gx = oracle.rxchfne_gam2_ic1(t)
    rg, rgx = oracle.ref_rxchfne_gam2_ic1(t)
    check_close(g, rg, what="RXCHFne_GAM2_IC1 GM2")
    check_close(gx, rgx, what="RXCHFne_GAM2_IC1 GM2ex")
    ne, npb = t["nebf"], t["npbf"]
    rng = np.random.default_rng(3)
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    DAa = synth.density(ne, 1.0, 4)
    for which in (0, 1, 2):
        n = npb if which == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        assert oracle.rxchfne_omg2(which, ne, npb, DA, DB, DP, g, A) == pytest.approx(
            oracle.ref_rxchfne_omg2(which, ne, npb, DA, DB, DP, g, B), rel=1e-12)
        check_close(A, B, what=f"RXCHFne omg2 {which}")
        A, B = F0.copy(order="F"), F0.copy(order="F")
        A2 = np.asfortranarray(rng.standard_normal((ne, ne)))
        B2 = A2.copy(order="F")
        assert oracle.rxcuhfne_omg2(which, ne, npb, DAa, DA, DB, DP, g, gx, A, A2) == pytest.approx(
            oracle.ref_rxcuhfne_omg2(which, ne, npb, DAa, DA, DB, DP, g, gx, B, B2), rel=1e-12)
        check_close(A, B, what=f"RXCUHFne omg2 {which}")
        if which == 0:
            check_close(A2, B2, what="RXCUHFne FE beta")
    gee = oracle.gam_ee(t)
    F0 = np.asfortranarray(rng.standard_normal((ne, ne)))
    A, B = F0.copy(order="F"), F0.copy(order="F")
    assert oracle.rxchfne_fae_gamee(ne, gee, DA, A) == pytest.approx(oracle.ref_rxchfne_fae_gamee(ne, gee, DA, B),
                                                                    rel=1e-12)
    check_close(A, B, what="RXCHFne_FAE_GAMee")


@needs_ref
def test_oracle_rxchfmult_vs_reference(tmp_path):
    """RXCHFmult: the two-basis e-e-p (three outputs) and e-e-e-p tuples,
    RXCHFmult_GAM2, _GAM3_INT, _GAM4_INT and the FAE/FP/FBE_OMG{2,3,4}
    contractions of the oracle against the reference's own routines
    (source/rxchfmult/)."""
    from _util import mult_tuples

    s = synth.mult(seed=50, e_shells=(0, 1, 2), be_shells=(0, 1, 2), p_shells=(0, 1, 2), be_centres=2)
    cs = synth.combined(s)
    for cls, n in ((5, 12), (6, 4)):
        idx = mult_tuples(s, cls, n, seed=cls)
        ref = oracle.ref_prim_list3(cs, cls, idx)
        got = oracle.prim_list3(cs, cls, idx)
        for k, what in enumerate(("ansE1", "ansS", "ansE2") if cls == 5 else ("ans",)):
            check_close(got[k], ref[k], what=f"cls{cls} {what}")
    t = synth.mult(seed=53, e_shells=(0, 1), be_shells=(0, 1), p_shells=(0, 0), n_e_centres=1, be_centres=1,
                   contract=(1,), ngtg=2)
    na, nb, npb = t["nebf"], t["nebfBE"], t["npbf"]
    assert nb <= na
    got = oracle.rxchfmult_gam2(t)
    ref = oracle.ref_rxchfmult_gam2(t)
    for g, r, what in zip(got, ref, ("INT_GAM2", "XCHF_GAM2", "XCHF_GAM2s")):
        check_close(g, r, what=what)
    g3 = oracle.rxchfmult_gam3_int(t)
    check_close(g3, oracle.ref_rxchfmult_gam3_int(t), what="INT_GAM3")
    rng = np.random.default_rng(8)
    gee = oracle.gam_ee(t)
    g4 = oracle.rxchfmult_gam4_int(na, nb, npb, got[2], gee)
    check_close(g4, oracle.ref_rxchfmult_gam4_int(na, nb, npb, got[2], gee, str(tmp_path)), what="INT_GAM4")
    assert np.max(np.abs(g4)) > 0.0
    DA, DB, DP = synth.density(na, 2.0, 1), synth.density(nb, 1.0, 2), synth.density(npb, 1.0, 3)
    for order, G in ((2, got[0]), (3, g3), (4, g4)):
        for which in (0, 1, 2):
            n = (na, npb, nb)[which]
            F0 = np.asfortranarray(rng.standard_normal((n, n)))
            A, B = F0.copy(order="F"), F0.copy(order="F")
            eo = oracle.rxchfmult_omg(order, which, na, nb, npb, DA, DB, DP, G, A)
            er = oracle.ref_rxchfmult_omg(order, which, na, nb, npb, DA, DB, DP, G, B)
            assert eo == pytest.approx(er, rel=1e-12, abs=1e-14)
            check_close(A - F0, B - F0, what=f"RXCHFmult omg{order} {which}")


@needs_ref
def test_oracle_rxchfmult_xchf_side_vs_reference(tmp_path):
    """RXCHFmult's XCHF-side builders over the special-electron basis:
    RXCHFmult_xcalc_GAM1_MD / RXCHFmult_XCHF_GAM3_MD tuples, RXCHFmult_GAM1_MD,
    RXCHFmult_GAM3_XCHF and RXCHFmult_GAM4_XCHF (GAM4_ICR over the special
    basis with the leading block of GAM_ee) against the reference."""
    s = synth.small(seed=55, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    for cls, n in ((8, 60), (7, 4)):
        nel = 1 if cls == 8 else 3
        idx = random_tuples(s, nel, n, seed=cls)
        ref = oracle.ref_prim_list3(s, cls, idx)
        got = oracle.prim_list3(s, cls, idx)
        check_close(got[0], ref[0], what=f"cls{cls} E")
        if cls == 8:
            check_close(got[1], ref[1], what="cls8 S")
    t = synth.small(seed=56, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 0), ngtg=2)
    g, gs = oracle.rxchfmult_gam1(t)
    rg, rgs = oracle.ref_rxchfmult_gam1(t, str(tmp_path))
    check_close(g, rg, what="RXCHFmult_GAM1_MD GAM_1")
    check_close(gs, rgs, what="RXCHFmult_GAM1_MD GAM_1s")
    x1, _ = oracle.gam1(t)
    assert np.max(np.abs(g - x1)) > 1e-6 * np.max(np.abs(x1))  # the one-geminal terms are really dropped
    u = synth.small(seed=57, n_e_centres=1, e_shells=(0, 1), contract=(1,), p_shells=(0,), ngtg=2)
    g3 = oracle.rxchfmult_gam3_xchf(u)
    check_close(g3, oracle.ref_rxchfmult_gam3_xchf(u), what="RXCHFmult_GAM3_XCHF")
    # GAM4_XCHF: special basis = the first nb functions of a larger regular basis
    reg = synth.small(seed=58, n_e_centres=1, e_shells=(0, 1, 0), contract=(1,), p_shells=(0,), ngtg=2)
    na, nb, npb = reg["nebf"], 4, reg["npbf"]
    gee = oracle.gam_ee(reg)
    rng = np.random.default_rng(2)
    g2s = rng.standard_normal(npb * npb * nb ** 4)
    eeb = np.ascontiguousarray(gee.reshape((na,) * 4)[:nb, :nb, :nb, :nb]).ravel()
    g4 = oracle.gam4_icr(nb, npb, g2s, eeb)
    check_close(g4, oracle.ref_rxchfmult_gam4_xchf(na, nb, npb, g2s, gee, str(tmp_path)), what="RXCHFmult_GAM4_XCHF")


@needs_ref
def test_oracle_rxchfmult_exchange_vs_reference():
    """RXCHFmult with exchange (one basis): RXCHFmult_GAM2_IC1ex,
    RXCHFmult_GAM3_IC1ex and the FAE/FP/FBE_OMG2ex and _OMG3ex contractions
    of the oracle against the reference (source/rxchfmult/)."""
    s = synth.small(seed=64, n_e_centres=1, e_shells=(0, 1), contract=(1,), p_shells=(0, 0), ngtg=2)
    got, ref = oracle.rxchfmult_gam2_ic1ex(s), oracle.ref_rxchfmult_gam2_ic1ex(s)
    for g, r, what in zip(got, ref, ("GM2_1ICR", "GM2_2ICR", "GM2_3ICR", "GM2sICR")):
        check_close(g, r, what=f"RXCHFmult_GAM2_IC1ex {what}")
    u = synth.small(seed=65, n_e_centres=1, e_shells=(0, 1), contract=(1,), p_shells=(0,), ngtg=2)
    g3, r3 = oracle.rxchfmult_gam3_ic1ex(u), oracle.ref_rxchfmult_gam3_ic1ex(u)
    for g, r, what in zip(g3, r3, ("INT_GAM3", "INT_GAM3ex1", "INT_GAM3ex2")):
        check_close(g, r, what=f"RXCHFmult_GAM3_IC1ex {what}")
    rng = np.random.default_rng(9)
    ne, npb = 3, 2
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    DA[0, 1] += 0.05  # unsymmetric densities tell the index orders apart
    for order in (2, 3):
        n = npb * npb * ne ** (2 * order)
        G, X1, X2 = (rng.standard_normal(n) for _ in range(3))
        for which in (0, 1, 2):
            m = npb if which == 1 else ne
            F0 = np.asfortranarray(rng.standard_normal((m, m)))
            A, B = F0.copy(order="F"), F0.copy(order="F")
            eo = oracle.rxchfmult_omgex(order, which, ne, npb, DA, DB, DP, G, X1, X2, A)
            er = oracle.ref_rxchfmult_omgex(order, which, ne, npb, DA, DB, DP, G, X1, X2, B)
            assert eo == pytest.approx(er, rel=1e-12, abs=1e-13)
            check_close(A - F0, B - F0, what=f"RXCHFmult omg{order}ex {which}")


# ----------------------------------------------------------------------
# NEO-XCHF RHF SCF side
# ----------------------------------------------------------------------
def _xsys(seed, nelec=4, **kw):
    s = synth.small(seed=seed, **kw)
    s["nelec"] = nelec
    return s


@needs_ref
def test_oracle_one_body_vs_reference(tmp_path):
    """class_nuc_rep, elec_ovlap / nuc_ovlap, calc_GAM_epcore and calc_GAM_ep
    (s/p/d on both particles) against the reference's routines."""
    s = _xsys(91, n_e_centres=2, e_shells=(0, 1, 2), contract=(1, 2), p_shells=(0, 1, 2))
    e, He, Se, Hp, Sp, Gep = oracle.ref_one_body(s, str(tmp_path))
    oHe, oSe, oHp, oSp = oracle.one_body(s)
    assert oracle.enuc(s) == pytest.approx(e, rel=1e-14)
    for o, r, nm in ((oHe, He, "GAM_ecore"), (oSe, Se, "S_e"), (oHp, Hp, "GAM_pcore"), (oSp, Sp, "S_p"),
                     (oracle.gam_ep(s), Gep, "GAM_ep")):
        check_close(o, r, what=nm)


@needs_ref
@pytest.mark.parametrize("n", [1, 5, 12])
def test_oracle_roothan_vs_reference(n):
    """ROOTHAN: eigenvalues and the density (electrons: the nocc lowest
    orbitals doubly occupied; nucleus: orbital NUCST singly)."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    S = A @ A.T / n + np.eye(n)
    F = rng.standard_normal((n, n))
    F = F + F.T
    for npart, ipart, nucst, k1, k2, pre in ((2 * max(1, n // 2), 1, 1, 1, max(1, n // 2), 2.0),
                                             (1, 2, min(2, n), min(2, n), min(2, n), 1.0)):
        D, Cm, ev = oracle.ref_roothan(S, F, npart, ipart, nucst)
        oC, oev, oD = oracle.roothan(S, F, k1, k2, pre)
        check_close(oev, ev, rtol=1e-12, atol_rel=1e-13, what="eigenvalues")
        check_close(oD, D, rtol=1e-10, atol_rel=1e-12, what="density")


@needs_ref
@pytest.mark.parametrize("nelec", [1, 2, 4])
def test_oracle_xchf_fock_vs_reference(tmp_path, nelec):
    """xchf1_fock (energy components, FE and FP after XCfock_correction) on
    the in-core path against the reference routine, with random densities."""
    s = _xsys(92, nelec=nelec, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,))
    ne, npb = s["nebf"], s["npbf"]
    ints = oracle.xchf_ints(s)
    DE, DP = synth.density(ne, float(nelec), 5), synth.density(npb, 1.0, 6)
    FE, FP, c = oracle.ref_xchf_fock(s, ints, DE, DP, str(tmp_path))
    oFE, oFP, oc = oracle.xchf_fock(ne, npb, nelec, ints, DE, DP)
    check_close(oc, c, what="components")
    check_close(oFE, FE, what="FE")
    check_close(oFP, FP, what="FP")


@pytest.mark.parametrize("name", ["tiny_s", "tiny_pnuc"])
def test_oracle_xcscf_vs_reference_run(name):
    """The oracle's SCF loop (ROOTHAN steps) converges to the energy the
    reference's xcscf (SOSCF steps) printed for the same system and starts
    from the same initial-guess components (tests/golden/xchf_scf.json)."""
    from _util import golden_scf, golden_system
    g = golden_scf()[name]
    s, guess = golden_system(g)
    r = oracle.xcscf(s, oracle.xchf_ints(s), ce_guess=guess)
    assert r["niter"] > 0
    assert abs(r["energy"] - g["energy"]) <= 1e-8, (r["energy"], g["energy"])
    names = ("E_TOTAL", "E_NUC", "E_ECORE", "E_PCORE", "E_EP", "E_EE", "E_GAM1", "E_GAM2", "E_GAM3", "E_GAM4",
             "S_TOTAL")
    for k, v in zip(names, r["comps_init"]):
        assert abs(v - g["initial"][k]) <= 1.5e-10 + 1e-12 * abs(v), k


def test_input_and_orbital_files():
    """The reference input reader reproduces the normalised coefficients and
    bohr coordinates examples/xchf/answer/XCHF.log prints (6 decimals), and
    the PUSQLF orbital file round-trips byte for byte."""
    from neorx import xchf
    d = os.path.join(HERE, "golden", "xchf_example")
    s = xchf.read_input(os.path.join(d, "exam_xchf.inp"))
    np.testing.assert_allclose(s["agebfcc"][:4], [0.441353, 0.693497, 0.664184, 0.287167], atol=5e-7)
    np.testing.assert_allclose(s["agebfcc"][8:], [0.163548, 0.277418, 0.315749, 0.147123], atol=5e-7)
    np.testing.assert_allclose(s["agnbfcc"], [2.077622], atol=5e-7)
    np.testing.assert_allclose(s["cat"][:, 2], [-1.75250354, 1.75250354], atol=5e-9)
    assert list(s["kpestr"]) == [1, 4, 5, 8, 9, 12] and list(s["kpeend"]) == [3, 4, 7, 8, 11, 12]
    C = xchf.read_mos(os.path.join(d, "guessCE.inp"), 6)
    out = os.path.join(HERE, "_build", "guess_roundtrip.dat")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    xchf.write_mos(out, C)
    assert open(out).read() == open(os.path.join(d, "guessCE.inp")).read()


@pytest.mark.parametrize("name", ["f_elec", "f_nuc"])
def test_oracle_f_tuples_vs_reference_golden(name):
    """f functions (l = 3): the oracle against the reference's xcalc_GAM{1,2,3}_MD
    values stored by tests/golden/make_f_tuples.py (the reference takes tens of
    seconds per f tuple, so they are not recomputed here)."""
    import json
    import os

    from _util import golden_system

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "f_tuples.json")))[name]
    s, _ = golden_system(g)
    for cls, ent in g["tuples"].items():
        idx = np.array(ent["idx"], dtype=np.int64)
        oe, os_ = oracle.prim_list(s, int(cls), idx)
        check_close(oe, np.array(ent["E"]), what=f"{name} cls{cls} E")
        if ent["S"] is not None:
            check_close(os_, np.array(ent["S"]), what=f"{name} cls{cls} S")


@needs_ref
def test_oracle_rxchf_vs_reference():
    """RXCHF (one special electron, source/rxchf/): the oracle's
    RXCHF_xcalc_GAM2_MD (class 9), RXCHF_GAM2_IC1 and RXCHF_FAE/FP/FBE_OMG2
    against the reference's own routines."""
    s = synth.small(seed=140, e_shells=(0, 1), p_shells=(0, 1))
    idx = random_tuples(s, 2, 6, seed=9)
    ref = oracle.ref_prim_list3(s, 9, idx)
    got = oracle.prim_list3(s, 9, idx)
    for k, what in enumerate(("ansE1", "ansS", "ansE2")):
        check_close(got[k], ref[k], what=f"RXCHF_xcalc_GAM2_MD {what}")
    t = synth.small(seed=141, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,), ngtg=2)
    ne, npb = t["nebf"], t["npbf"]
    g = oracle.rxchf_gam2_ic1(t)
    rg = oracle.ref_rxchf_gam2_ic1(t)
    for a, b, what in zip(g, rg, ("GM2_1ICR", "GM2_2ICR", "GM2sICR")):
        check_close(a, b, what=f"RXCHF_GAM2_IC1 {what}")
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    rng = np.random.default_rng(3)
    for w in range(3):
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        S0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B, C2, D = F0.copy(order="F"), F0.copy(order="F"), S0.copy(order="F"), S0.copy(order="F")
        eo = oracle.rxchf_omg2(w, ne, npb, DA, DB, DP, *g, A, C2)
        er = oracle.ref_rxchf_omg2(w, ne, npb, DA, DB, DP, *rg, B, D)
        check_close(A - F0, B - F0, what=f"RXCHF OMG2 {w} F")
        check_close(C2 - S0, D - S0, what=f"RXCHF OMG2 {w} S")
        assert eo[0] == pytest.approx(er[0], rel=1e-12) and eo[1] == pytest.approx(er[1], rel=1e-12)


@needs_ref
def test_oracle_rxchf3_vs_reference(tmp_path):
    """RXCHF (one special electron): the oracle's RXCHF_xcalc_GAM3_MD (class
    10), RXCHF_GAM3_IC1, RXCHF_FAE/FP/FBE_OMG3 and RXCHF_FP/FBE_OMG1 against
    the reference's own routines.  The contractions run over random tensors
    (every index permutation of the formulas touches distinct values), OMG1
    over GAM_1 files written for the reference to read."""
    s = synth.small(seed=142, e_shells=(0, 1), p_shells=(0, 1))
    idx = random_tuples(s, 3, 8, seed=10)
    ref = oracle.ref_prim_list3(s, 10, idx)
    got = oracle.prim_list3(s, 10, idx)
    check_close(got[0], ref[0], what="RXCHF_xcalc_GAM3_MD ans1")
    check_close(got[1], ref[1], what="RXCHF_xcalc_GAM3_MD ans2")
    t = synth.small(seed=143, n_e_centres=2, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    g = oracle.rxchf_gam3_ic1(t)
    rg = oracle.ref_rxchf_gam3_ic1(t)
    check_close(g[0], rg[0], what="RXCHF_GAM3_IC1 GM3_1ICR")
    check_close(g[1], rg[1], what="RXCHF_GAM3_IC1 GM3_2ICR")
    ne, npb = 3, 2
    rng = np.random.default_rng(4)
    G1, G2 = rng.standard_normal(ne ** 6 * npb * npb), rng.standard_normal(ne ** 6 * npb * npb)
    DA, DB, DP = synth.density(ne, 2.0, 5), synth.density(ne, 1.0, 6), synth.density(npb, 1.0, 7)
    for w in range(3):
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        eo = oracle.rxchf_omg3(w, ne, npb, DA, DB, DP, G1, G2, A)
        er = oracle.ref_rxchf_omg3(w, ne, npb, DA, DB, DP, G1, G2, B)
        check_close(A - F0, B - F0, what=f"RXCHF OMG3 {w} F")
        assert eo == pytest.approx(er, rel=1e-12)
    H1, H1s = rng.standard_normal(ne * ne * npb * npb), rng.standard_normal(ne * ne * npb * npb)
    for w in (1, 2):
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        SA, SB = np.asfortranarray(rng.standard_normal((n, n))), np.zeros((n, n), order="F")
        eo = oracle.rxchf_omg1(w, ne, npb, DB, DP, H1, H1s, A, SA)
        er = oracle.ref_rxchf_omg1(w, ne, npb, DB, DP, H1, H1s, B, SB, str(tmp_path))
        check_close(A - F0, B - F0, what=f"RXCHF OMG1 {w} F")
        check_close(SA, SB, what=f"RXCHF OMG1 {w} S")
        assert eo[0] == pytest.approx(er[0], rel=1e-12) and eo[1] == pytest.approx(er[1], rel=1e-12)


@needs_ref
def test_oracle_rxcuhf_vs_reference():
    """RXCUHF (open-shell regular electrons, one special electron): the
    oracle's RXCUHF_FE/FP/FBE_OMG2 and _OMG3 against the reference's routines
    over random tensors and densities (alpha and beta densities differ)."""
    ne, npb = 3, 2
    rng = np.random.default_rng(8)
    Dal, Dbt, DB, DP = (synth.density(n, t, sd) for n, t, sd in ((ne, 2.0, 11), (ne, 1.0, 12), (ne, 1.0, 13),
                                                                  (npb, 1.0, 14)))
    Dto = Dal + Dbt
    G2a = [rng.standard_normal(ne ** 4 * npb * npb) for _ in range(3)]
    G3a = [rng.standard_normal(ne ** 6 * npb * npb) for _ in range(2)]
    for w in range(3):
        n = npb if w == 1 else ne
        F0, S0, B0 = (np.asfortranarray(rng.standard_normal((n, n))) for _ in range(3))
        A, SA, BA = F0.copy(order="F"), S0.copy(order="F"), B0.copy(order="F")
        R, SR, BR = F0.copy(order="F"), S0.copy(order="F"), B0.copy(order="F")
        eo = oracle.rxcuhf_omg2(w, ne, npb, Dal, Dto, DB, DP, *G2a, A, SA, BA)
        er = oracle.ref_rxcuhf_omg2(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G2a, R, SR, BR)
        check_close(A - F0, R - F0, what=f"RXCUHF OMG2 {w} F")
        check_close(SA - S0, SR - S0, what=f"RXCUHF OMG2 {w} S")
        if w == 0:
            check_close(BA - B0, BR - B0, what="RXCUHF OMG2 FAbetE")
        assert eo[0] == pytest.approx(er[0], rel=1e-12) and eo[1] == pytest.approx(er[1], rel=1e-12)
        F0, B0 = (np.asfortranarray(rng.standard_normal((n, n))) for _ in range(2))
        A, BA, R, BR = F0.copy(order="F"), B0.copy(order="F"), F0.copy(order="F"), B0.copy(order="F")
        eo = oracle.rxcuhf_omg3(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G3a, A, BA)
        er = oracle.ref_rxcuhf_omg3(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G3a, R, BR)
        check_close(A - F0, R - F0, what=f"RXCUHF OMG3 {w} F")
        if w == 0:
            check_close(BA - B0, BR - B0, what="RXCUHF OMG3 FAbetE")
        assert eo == pytest.approx(er, rel=1e-12)
