"""The device integral engine (nrx_engine.h) compiled for the host, checked
against the oracle before any GPU run.  The same source is what the HIP
kernels execute."""
import numpy as np
import pytest

import oracle
from neorx import synth
from _util import check_close, engine_prim_list, random_tuples


@pytest.mark.parametrize("cls,n", [(1, 400), (2, 200), (3, 60)])
def test_engine_tuples_vs_oracle(cls, n):
    s = synth.small(seed=40 + cls, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    idx = random_tuples(s, cls, n, seed=7 * cls)
    ee, es = engine_prim_list(s, cls, idx)
    oe, os_ = oracle.prim_list(s, cls, idx)
    check_close(ee, oe, what=f"engine vs oracle cls{cls} E")
    if cls < 3:
        check_close(es, os_, what=f"engine vs oracle cls{cls} S")


def test_engine_config2_sample():
    s = synth.config2()
    for cls, n in [(1, 200), (2, 100), (3, 30)]:
        idx = random_tuples(s, cls, n, seed=cls)
        ee, _ = engine_prim_list(s, cls, idx)
        oe, _ = oracle.prim_list(s, cls, idx)
        check_close(ee, oe, what=f"config2 cls{cls}")


def test_engine_large_T_branch():
    """Far-apart centres push the Boys argument past 20 (asymptotic branch)."""
    s = synth.small(seed=9, e_shells=(0, 1), p_shells=(0, 1))
    s["elcbfc"] = np.asfortranarray(s["elcbfc"] * 4.0)
    s["elcex"] = s["elcex"] * 3.0
    idx = random_tuples(s, 2, 120, seed=3)
    ee, _ = engine_prim_list(s, 2, idx)
    oe, _ = oracle.prim_list(s, 2, idx)
    check_close(ee, oe, what="large-T")


def test_engine_boys_vs_reference_semantics():
    """Device Boys function: identical to the reference's asymptotic branch
    for T > 20; for T <= 20 within 4e-15 of the exact function (a 40-digit
    series) and within 1e-11 of the reference's series, whose absolute 1e-12
    cut leaves up to ~3e-13 relative error at high order (F_12(0.256))."""
    import ctypes as C
    import math

    from _util import engine_host

    E = engine_host()
    F = (C.c_double * 40)()
    for T in np.concatenate([np.linspace(0.0, 20.0, 157), [20.0, 20.0000001, 23.7, 41.0, 133.0]]):
        M = 17
        E.eng_boys(C.c_int64(M), C.c_double(T), F)
        for m in range(M + 1):
            ref = oracle.boys(m, T)  # the oracle's per-order restatement of iboys2
            if T > 20:
                asym = 0.88622692545275801365 / math.sqrt(T)
                for l in range(1, m + 1):
                    asym = asym * (l - 0.5) / T
                assert F[m] == asym
            else:
                assert abs(F[m] - ref) <= 1e-11 * ref, (m, T, F[m], ref)
    from decimal import Decimal, getcontext

    getcontext().prec = 40
    for T in [0.0, 0.2564102564102564, 3.3, 11.9, 19.97]:
        E.eng_boys(C.c_int64(16), C.c_double(T), F)
        for m in (0, 5, 12, 16):
            Td = Decimal(T)
            term = Decimal(1) / Decimal(2 * m + 1)
            acc = term
            for k in range(1, 400):
                term = term * 2 * Td / Decimal(2 * m + 2 * k + 1)
                acc += term
            exact = float(acc * (-Td).exp())
            assert abs(F[m] - exact) <= 4e-15 * exact, (m, T, F[m], exact)


def test_engine_rxchfne_tuples_vs_oracle():
    """RXCHFne_xcalc_GAM2_MD (two outputs) in the host build of the engine
    against the oracle, both capacity tiers."""
    s = synth.small(seed=47, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    idx = random_tuples(s, 2, 200, seed=8)
    oe, os_ = oracle.prim_list(s, 4, idx)
    for tier in (0, 1):
        ee, es = engine_prim_list(s, 4, idx, tier)
        check_close(ee, oe, what=f"RXCHF-ne ansE1 tier {tier}")
        check_close(es, os_, what=f"RXCHF-ne ansE2 tier {tier}")


def test_engine_rxchfmult_tuples_vs_oracle():
    """RXCHFmult_GAM2_MD (three outputs) and RXCHFmult_INT_GAM3_MD in the host
    build of the engine against the oracle, both capacity tiers."""
    from _util import engine_prim_list3, mult_tuples

    s = synth.mult(seed=54, e_shells=(0, 1, 2), be_shells=(0, 1, 2), p_shells=(0, 1, 2), be_centres=2)
    cs = synth.combined(s)
    for cls, n in ((5, 120), (6, 40)):
        idx = mult_tuples(s, cls, n, seed=10 + cls)
        ref = oracle.prim_list3(cs, cls, idx)
        for tier in (0, 1):
            got = engine_prim_list3(cs, cls, idx, tier)
            for k, what in enumerate(("ansE1", "ansS", "ansE2") if cls == 5 else ("ans",)):
                check_close(got[k], ref[k], what=f"cls{cls} {what} tier {tier}")


def test_engine_rxchfmult_xchf_side_vs_oracle():
    """RXCHFmult_xcalc_GAM1_MD (class 8) and RXCHFmult_XCHF_GAM3_MD (class 7)
    in the host engine against the oracle; class 7 keeps the labels of
    xcalc_GAM3_MD, so it is compared after averaging each tuple over the six
    permutations of its electrons."""
    from itertools import permutations

    from _util import engine_prim_list3

    s = synth.small(seed=59, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    idx = random_tuples(s, 1, 300, seed=21)
    ref = oracle.prim_list3(s, 8, idx)
    for tier in (0, 1):
        got = engine_prim_list3(s, 8, idx, tier)
        check_close(got[0], ref[0], what=f"cls8 E tier {tier}")
        check_close(got[1], ref[1], what=f"cls8 S tier {tier}")
    base = random_tuples(s, 3, 20, seed=22)
    rows = []
    for x in base:
        for p in permutations(range(3)):
            rows.append([x[p[0]], x[p[1]], x[p[2]], x[3], x[4 + p[0]], x[4 + p[1]], x[4 + p[2]], x[7]])
    idx = np.array(rows, dtype=np.int64)
    ref = oracle.prim_list3(s, 7, idx)[0].reshape(-1, 6).mean(axis=1)
    for tier in (0, 1):
        got = engine_prim_list3(s, 7, idx, tier)[0].reshape(-1, 6).mean(axis=1)
        check_close(got, ref, what=f"cls7 six-permutation average tier {tier}")


def _lone_am_tuples(s, cls):
    """Tuples of class cls whose angular momentum sits on one particle: every
    particle s-type except one, whose bra and ket are d-type along one axis
    (or d and p).  These put one particle's polynomial at degree 4 with the
    rest at 0, the case where a one-particle kinetic operator needs more
    capacity than the summed degrees suggest."""
    am, pm = np.asarray(s["elcam"]), np.asarray(s["nucam"])
    ne = {1: 1, 2: 2, 3: 3, 4: 2, 7: 3, 8: 1}[cls]
    e_s = int(np.flatnonzero(am.sum(1) == 0)[0])
    p_s = int(np.flatnonzero(pm.sum(1) == 0)[0])
    e_d = [int(i) for i in np.flatnonzero(am.max(1) == 2)[:3]]
    e_p = [int(i) for i in np.flatnonzero(am.sum(1) == 1)[:3]]
    p_d = [int(i) for i in np.flatnonzero(pm.max(1) == 2)[:3]]
    rows = []
    for j in range(ne + 1):
        heavy = p_d if j == ne else e_d
        other = e_p if j < ne else [int(i) for i in np.flatnonzero(pm.sum(1) == 1)[:3]]
        for a in heavy:
            for b in heavy + other:
                bra = [e_s] * ne + [p_s]
                ket = [e_s] * ne + [p_s]
                bra[j], ket[j] = a, b
                rows.append(bra + ket)
    return np.asarray(rows, dtype=np.int64)


@pytest.mark.parametrize("cls", [1, 2, 3, 4, 7, 8])
def test_engine_tiers_lone_angular_momentum(cls):
    """Every capacity tier gives the general tier's values (and the
    oracle's) on tuples with all angular momentum on one particle, and on a
    random sample; regression for a tier that lacked the +2 of the e-e-e-p
    tuple's one-particle kinetic operator on e3."""
    from _util import engine_prim_list3

    s = synth.config2()
    ne = {1: 1, 2: 2, 3: 3, 4: 2, 7: 3, 8: 1}[cls]
    idx = np.concatenate([_lone_am_tuples(s, cls), random_tuples(s, ne, 150, seed=90 + cls)])
    if cls in (7, 8):
        t0 = engine_prim_list3(s, cls, idx, 0)[0]
        t1 = engine_prim_list3(s, cls, idx, 1)[0]
        ref = oracle.prim_list3(s, cls, idx)[0] if cls == 8 else None
    else:
        t0 = engine_prim_list(s, cls, idx, 0)[0]
        t1 = engine_prim_list(s, cls, idx, 1)[0]
        ref = oracle.prim_list(s, cls, idx)[0]
    check_close(t0, t1, rtol=1e-13, atol_rel=1e-14, what=f"cls{cls} device tiers vs general tier")
    if ref is not None:
        check_close(t0, ref, what=f"cls{cls} device tiers vs oracle")


@pytest.mark.parametrize("cls,n", [(1, 300), (2, 120), (3, 40), (4, 80), (7, 20), (8, 200)])
def test_engine_f_tuples_vs_oracle(cls, n):
    """The l <= 3 build of the engine (the wide bin of the device kernels) on
    s/p/d/f electrons and an s/f nucleus; the l <= 2 build's tier choice sends
    every tuple with an f power to that bin.  Class 7 equals the reference
    after the average over the six permutations of its electron pairs."""
    from itertools import permutations

    s = synth.small(seed=130 + cls, n_e_centres=2, e_shells=(0, 2, 3), p_shells=(0, 3))
    idx = random_tuples(s, {4: 2, 7: 3, 8: 1}.get(cls, cls), n, seed=cls)
    if cls == 7:
        idx = np.array([[x[p[0]], x[p[1]], x[p[2]], x[3], x[4 + p[0]], x[4 + p[1]], x[4 + p[2]], x[7]]
                        for x in idx for p in permutations(range(3))], dtype=np.int64)
    ee, es = engine_prim_list(s, cls, idx, lmax=3)
    oe, os_ = oracle.prim_list(s, cls, idx)
    if cls == 7:
        ee, oe = ee.reshape(-1, 6).mean(axis=1), oe.reshape(-1, 6).mean(axis=1)
    # absolute floor 1e-11 of the batch maximum: the oracle keeps the
    # reference's truncated Boys series (~5e-12 relative at the orders f
    # functions reach), the engine the exact function; cancelling terms show it
    check_close(ee, oe, atol_rel=1e-11, what=f"l<=3 engine vs oracle cls{cls} out0")
    if cls not in (3, 7):
        check_close(es, os_, atol_rel=1e-11, what=f"l<=3 engine vs oracle cls{cls} out1")


def test_engine_f_rxchfmult_vs_oracle():
    """Two-basis classes 5 and 6 with f functions, l <= 3 build."""
    from _util import engine_prim_list3, mult_tuples

    s = synth.mult(seed=137, e_shells=(0, 3), be_shells=(0, 1, 3), p_shells=(0, 1), be_centres=1)
    cs = synth.combined(s)
    for cls, n in ((5, 60), (6, 24)):
        idx = mult_tuples(s, cls, n, seed=cls)
        got = engine_prim_list3(cs, cls, idx, 0, lmax=3)
        ref = oracle.prim_list3(cs, cls, idx)
        for k in range(3 if cls == 5 else 1):
            check_close(got[k], ref[k], atol_rel=1e-11, what=f"l<=3 engine cls{cls} out{k}")


@pytest.mark.parametrize("lmax,shells", [(2, (0, 1, 2)), (3, (0, 3))])
def test_engine_rxchf_tuples_vs_oracle(lmax, shells):
    """RXCHF_xcalc_GAM2_MD (class 9: ansE1, ansS, ansE2), every capacity tier
    (tier 0) and the general one (tier 1); the l <= 3 build with f
    functions."""
    from _util import engine_prim_list3

    s = synth.small(seed=142 + lmax, e_shells=shells, p_shells=(0, 1, 2))
    idx = random_tuples(s, 2, 150, seed=5)
    ref = oracle.prim_list3(s, 9, idx)
    for tier in (0, 1):
        got = engine_prim_list3(s, 9, idx, tier, lmax=lmax)
        for k in range(3):
            check_close(got[k], ref[k], atol_rel=1e-12 if lmax == 2 else 1e-11,
                        what=f"class 9 out{k} tier {tier} lmax {lmax}")



@pytest.mark.parametrize("lmax,shells", [(2, (0, 1, 2)), (3, (0, 3))])
def test_engine_rxchf3_tuples_vs_oracle(lmax, shells):
    """RXCHF_xcalc_GAM3_MD (class 10: ans1, ans2), every capacity tier (tier
    0) and the general one (tier 1); the l <= 3 build with f functions."""
    from _util import engine_prim_list3

    s = synth.small(seed=150 + lmax, e_shells=shells, p_shells=(0, 1))
    idx = random_tuples(s, 3, 60 if lmax == 2 else 24, seed=6)
    ref = oracle.prim_list3(s, 10, idx)
    for tier in (0, 1):
        got = engine_prim_list3(s, 10, idx, tier, lmax=lmax)
        for k in range(2):
            check_close(got[k], ref[k], atol_rel=1e-12 if lmax == 2 else 1e-11,
                        what=f"class 10 out{k} tier {tier} lmax {lmax}")
