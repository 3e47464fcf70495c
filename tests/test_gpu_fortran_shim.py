"""The Fortran drop-in subroutines (neo-rxchf_amd/fortran/neorx_shim.f90)
called from a Fortran program with the reference's argument lists, checked
against the oracle.  The GPU test runs the program; the CPU test checks that
it links and resolves every library."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from neorx import synth
from _util import check_close

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_build", "drive_shim")


def test_driver_links():
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True).stdout
    assert "libneorx_fshim.so" in out and "libneorx_hip.so" in out
    assert "not found" not in out


def _write_input(path, s, DE, DP):
    with open(path, "wb") as f:
        np.array([s["nebf"], s["npebf"], s["npbf"], s["nat"], s["ngtg"]], dtype=np.int64).tofile(f)
        np.array([s["pmass"]], dtype=np.float64).tofile(f)
        for k in ("cat", "zan", "bcoef", "gamma"):
            np.ascontiguousarray(s[k], dtype=np.float64).ravel().tofile(f)
        for k in ("kpestr", "kpeend", "ampeb2c"):
            np.ascontiguousarray(s[k], dtype=np.int64).tofile(f)
        for k in ("agebfcc", "agnbfcc", "elcex", "nucex"):
            np.ascontiguousarray(s[k], dtype=np.float64).tofile(f)
        for k in ("elcam", "nucam"):
            np.asarray(s[k], dtype=np.int64).ravel(order="F").tofile(f)
        for k in ("elcbfc", "nucbfc"):
            np.asarray(s[k], dtype=np.float64).ravel(order="F").tofile(f)
        np.asarray(DE).ravel(order="F").tofile(f)
        np.asarray(DP).ravel(order="F").tofile(f)


@pytest.mark.gpu
@pytest.mark.parametrize("recl", [8, 32])
def test_fortran_dropin_matches_oracle(recl):
    """recl: .ufm record length (NEORX_UFM_RECL): 8 bytes as flang/gfortran
    write RECL=8, 32 as an ifort-built driver does (each double followed by 24
    zero bytes); the driver's GAM4_ICR reads back the GAM_ee.ufm it wrote."""
    s = synth.small(seed=81, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,), ngtg=2)
    ne, npb = s["nebf"], s["npbf"]
    sub = dict(s)
    # keep the nebf^8 in-core GAM_4 small: first 3 contracted functions
    keep = 3
    last = s["kpeend"][keep - 1]
    for k in ("elcex", "agebfcc"):
        sub[k] = s[k][:last].copy()
    for k in ("elcam", "elcbfc"):
        sub[k] = np.asfortranarray(s[k][:last])
    sub["ampeb2c"] = s["ampeb2c"][:last].copy()
    sub["kpestr"], sub["kpeend"] = s["kpestr"][:keep].copy(), s["kpeend"][:keep].copy()
    sub["nebf"], sub["npebf"] = keep, int(last)
    s = sub
    ne, npb = s["nebf"], s["npbf"]
    DE, DP = synth.density(ne, 2.0, 1), synth.density(npb, 1.0, 2)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        _write_input(fin, s, DE, DP)
        print(f"case fortran drop-in driver recl={recl}", flush=True)
        env = dict(os.environ, NEORX_UFM_RECL=str(recl))
        r = subprocess.run([DRIVER, fin, fout], cwd=td, capture_output=True, text=True, timeout=240, env=env)
        assert r.returncode == 0, r.stderr
        out = np.fromfile(fout, dtype=np.float64)
        stride = recl // 8
        raw1 = np.fromfile(os.path.join(td, "GAM_1.ufm"), dtype=np.float64).reshape(-1, stride)
        rawe = np.fromfile(os.path.join(td, "GAM_ee.ufm"), dtype=np.float64).reshape(-1, stride)
        assert rawe.shape[0] == ne ** 4 and not np.any(rawe[:, 1:]) and not np.any(raw1[:, 1:])
        gam1, gee = raw1[:, 0].copy(), rawe[:, 0].copy()
    ng2, ng3, ng4 = ne ** 4 * npb ** 2, ne ** 6 * npb ** 2, ne ** 8 * npb ** 2
    sizes = [ng2, ng2, ne * ne, ne * ne, npb * npb, npb * npb, 1, 1, ng3, ne * ne, npb * npb, 1, ng4, ne * ne,
             npb * npb, 1,
             ne * ne, npb * npb, ne * ne, npb * npb, 1, 1,  # GAM2_DSCF: FE, FP, SE2, SP2, E, S
             ne * ne, npb * npb, 1,  # GAM3_DSCF
             ne * ne, npb * npb, 1,  # GAM4_DSCF
             ng4,  # OMG4_ICR
             ne * ne, ne * ne, 1, 1,  # FAE_OMG2: FAE, SAE2, E, S
             npb * npb, npb * npb, 1, 1,  # UFP_OMG2
             ne * ne, 1, npb * npb, 1,  # FAE_OMG3, UFP_OMG3
             ne * ne, 1, npb * npb, 1,  # FAE_OMG4, UFP_OMG4
             ng2, ng2,  # RXCHFne_GAM2_IC1
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # RXCHFne_FAE/FP/FBE_OMG2
             ne * ne, ne * ne, 1,  # RXCUHFne_FE_OMG2
             ng2, ng2, ng2,  # RXCHFmult_GAM2: INT_GAM2, XCHF_GAM2, XCHF_GAM2s
             ng3, ng4,  # RXCHFmult_GAM3_INT, RXCHFmult_GAM4_INT
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # RXCHFmult_FAE/FP/FBE_OMG2
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # ... _OMG3
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # ... _OMG4
             ne * ne * npb * npb, ne * ne * npb * npb,  # GAM_1 of GAM1_OMP_MD, of RXCHFmult_GAM1_MD
             ng3, ng4,  # RXCHFmult_GAM3_XCHF, RXCHFmult_GAM4_XCHF (in the driver gm2s holds XCHF_GAM2s)
             ng2, ng2, ng2, ng2,  # RXCHFmult_GAM2_IC1ex
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # RXCHFmult_FAE/FP/FBE_OMG2ex
             ng3, ng3, ng3,  # RXCHFmult_GAM3_IC1ex
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # RXCHFmult_FAE/FP/FBE_OMG3ex
             ng2, ng2, ng2,  # RXCHF_GAM2_IC1: GM2_1ICR, GM2_2ICR, GM2sICR
             ne * ne, ne * ne, 1, 1, npb * npb, npb * npb, 1, 1, ne * ne, ne * ne, 1, 1,  # RXCHF_FAE/FP/FBE_OMG2
             ng3, ng3,  # RXCHF_GAM3_IC1: GM3_1ICR, GM3_2ICR
             ne * ne, 1, npb * npb, 1, ne * ne, 1,  # RXCHF_FAE/FP/FBE_OMG3
             npb * npb, npb * npb, 1, 1, ne * ne, ne * ne, 1, 1,  # RXCHF_FP/FBE_OMG1
             ne * ne, ne * ne, ne * ne, 1, 1, npb * npb, npb * npb, 1, 1, ne * ne, ne * ne, 1, 1,  # RXCUHF OMG2
             ne * ne, ne * ne, 1, npb * npb, 1, ne * ne, 1]  # RXCUHF OMG3
    assert out.size == sum(sizes)
    parts = np.split(out, np.cumsum(sizes)[:-1])
    og2, og2s = oracle.gam2_ic1(s)
    check_close(parts[0], og2, what="GAM2_IC1")
    check_close(parts[1], og2s, what="GAM2_IC1 S")
    FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
    SE, SP, E, S = oracle.eg2ic1(ne, npb, DE, DP, og2, og2s, FE, FP)
    check_close(parts[2], FE.ravel(order="F"), what="EG2IC1 FE")
    check_close(parts[3], SE.ravel(order="F"), what="EG2IC1 SE")
    check_close(parts[4], FP.ravel(order="F"), what="EG2IC1 FP")
    assert parts[6][0] == pytest.approx(E, rel=1e-11) and parts[7][0] == pytest.approx(S, rel=1e-11)
    og3 = oracle.gam3_ic1(s)
    check_close(parts[8], og3, what="GAM3_IC1")
    FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
    E3 = oracle.eg3ic1(ne, npb, DE, DP, og3, FE, FP)
    check_close(parts[9], FE.ravel(order="F"), what="EG3IC1 FE")
    assert parts[11][0] == pytest.approx(E3, rel=1e-11)
    oee = oracle.gam_ee(s)
    check_close(gee, oee, what="calc_GAM_ee file")
    og4 = oracle.gam4_icr(ne, npb, og2s, oee)
    check_close(parts[12], og4, what="GAM4_ICR")
    FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
    E4 = oracle.eg4icr(ne, npb, DE, DP, og4, FE, FP)
    check_close(parts[13], FE.ravel(order="F"), what="EG4ICR FE")
    assert parts[15][0] == pytest.approx(E4, rel=1e-11)
    og1, og1s = oracle.gam1(s)
    check_close(parts[79], og1, what="GAM_1 file of GAM1_OMP_MD")
    check_close(gam1, parts[80], what="GAM_1 file of RXCHFmult_GAM1_MD")
    # direct-SCF routines against the oracle's restatement of GAM*_DSCF
    F = [np.zeros((n, n), order="F") for n in (ne, npb, ne, npb)]
    E2, S2 = oracle.gam2_dscf(s, DE, DP, *F)
    for k, nm in enumerate(("FE", "FP", "SE2", "SP2")):
        check_close(parts[16 + k], F[k].ravel(order="F"), what=f"GAM2_DSCF {nm}")
    assert parts[20][0] == pytest.approx(E2, rel=1e-11) and parts[21][0] == pytest.approx(S2, rel=1e-11)
    F = [np.zeros((n, n), order="F") for n in (ne, npb)]
    E3d = oracle.gam3_dscf(s, DE, DP, *F)
    check_close(parts[22], F[0].ravel(order="F"), what="GAM3_DSCF FE")
    check_close(parts[23], F[1].ravel(order="F"), what="GAM3_DSCF FP")
    assert parts[24][0] == pytest.approx(E3d, rel=1e-11)
    F = [np.zeros((n, n), order="F") for n in (ne, npb)]
    E4d = oracle.gam4_dscf(ne, npb, og2s, oee, DE, DP, *F)
    check_close(parts[25], F[0].ravel(order="F"), what="GAM4_DSCF FE")
    check_close(parts[26], F[1].ravel(order="F"), what="GAM4_DSCF FP")
    assert parts[27][0] == pytest.approx(E4d, rel=1e-11)
    # XCUHF pieces, NAE=2, NBE=1, DAE = 2/3 DE, DBE = 1/3 DE
    og4u = oracle.omg4_icr(ne, npb, og2s, oee)
    check_close(parts[28], og4u, what="OMG4_ICR")
    DAE, DBE = DE * (2.0 / 3.0), DE * (1.0 / 3.0)
    a = (2, 1, ne, npb, DAE, DBE, DP)
    F, S2 = np.zeros((ne, ne), order="F"), np.zeros((ne, ne), order="F")
    e, sv = oracle.fae_omg2(*a, og2, og2s, F, S2)
    check_close(parts[29], F.ravel(order="F"), what="FAE_OMG2")
    check_close(parts[30], S2.ravel(order="F"), what="FAE_OMG2 SAE2")
    assert parts[31][0] == pytest.approx(e, rel=1e-11) and parts[32][0] == pytest.approx(sv, rel=1e-11)
    F, S2 = np.zeros((npb, npb), order="F"), np.zeros((npb, npb), order="F")
    e, sv = oracle.ufp_omg2(*a, og2, og2s, F, S2)
    check_close(parts[33], F.ravel(order="F"), what="UFP_OMG2")
    check_close(parts[34], S2.ravel(order="F"), what="UFP_OMG2 SP2")
    assert parts[35][0] == pytest.approx(e, rel=1e-11) and parts[36][0] == pytest.approx(sv, rel=1e-11)
    for k, (fn, G, n) in enumerate(((oracle.fae_omg3, og3, ne), (oracle.ufp_omg3, og3, npb),
                                    (oracle.fae_omg4, og4u, ne), (oracle.ufp_omg4, og4u, npb))):
        F = np.zeros((n, n), order="F")
        e = fn(*a, G, F)
        check_close(parts[37 + 2 * k], F.ravel(order="F"), what=fn.__name__)
        assert parts[38 + 2 * k][0] == pytest.approx(e, rel=1e-11, abs=1e-14), fn.__name__
    # RXCHF-ne pieces: regular density DAE = 2/3 DE, special DBE = 1/3 DE
    rg, rgx = oracle.rxchfne_gam2_ic1(s)
    check_close(parts[45], rg, what="RXCHFne_GAM2_IC1 GM2")
    check_close(parts[46], rgx, what="RXCHFne_GAM2_IC1 GM2ex")
    for k in range(3):
        n = npb if k == 1 else ne
        F = np.zeros((n, n), order="F")
        e = oracle.rxchfne_omg2(k, ne, npb, DAE, DBE, DP, rg, F)
        check_close(parts[47 + 2 * k], F.ravel(order="F"), what=f"RXCHFne omg2 {k}")
        assert parts[48 + 2 * k][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    Fa, Fb = np.zeros((ne, ne), order="F"), np.zeros((ne, ne), order="F")
    e = oracle.rxcuhfne_omg2(0, ne, npb, DBE, DE, DBE, DP, rg, rgx, Fa, Fb)  # DAalpE, DAtotE as the driver passes
    check_close(parts[53], Fa.ravel(order="F"), what="RXCUHFne_FE_OMG2 alpha")
    check_close(parts[54], Fb.ravel(order="F"), what="RXCUHFne_FE_OMG2 beta")
    assert parts[55][0] == pytest.approx(e, rel=1e-11, abs=1e-14)

    # RXCHFmult pieces, special-electron basis = regular basis
    m = dict(s)
    m["nebfBE"], m["npebfBE"] = s["nebf"], s["npebf"]
    for k in synth.BE_KEYS:
        m[k + "_be"] = s[k]
    mi, mx, mxs = oracle.rxchfmult_gam2(m)
    check_close(parts[56], mi, what="RXCHFmult_GAM2 INT_GAM2")
    check_close(parts[57], mx, what="RXCHFmult_GAM2 XCHF_GAM2")
    check_close(parts[58], mxs, what="RXCHFmult_GAM2 XCHF_GAM2s")
    m3 = oracle.rxchfmult_gam3_int(m)
    check_close(parts[59], m3, what="RXCHFmult_GAM3_INT")
    m4 = oracle.rxchfmult_gam4_int(ne, ne, npb, mxs, oee)
    check_close(parts[60], m4, what="RXCHFmult_GAM4_INT")
    for order, G in ((2, mi), (3, m3), (4, m4)):
        for which in range(3):
            n = npb if which == 1 else ne
            F = np.zeros((n, n), order="F")
            e = oracle.rxchfmult_omg(order, which, ne, ne, npb, DAE, DBE, DP, G, F)
            at = 61 + 6 * (order - 2) + 2 * which
            check_close(parts[at], F.ravel(order="F"), what=f"RXCHFmult omg{order} {which}")
            assert parts[at + 1][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    check_close(parts[80], oracle.rxchfmult_gam1(s)[0], what="RXCHFmult_GAM1_MD")
    check_close(parts[81], oracle.rxchfmult_gam3_xchf(s), what="RXCHFmult_GAM3_XCHF")
    check_close(parts[82], oracle.gam4_icr(ne, npb, mxs, oee), what="RXCHFmult_GAM4_XCHF")
    # RXCHFmult with exchange (one basis)
    x2 = oracle.rxchfmult_gam2_ic1ex(s)
    for k, what in enumerate(("GM2_1ICR", "GM2_2ICR", "GM2_3ICR", "GM2sICR")):
        check_close(parts[83 + k], x2[k], what=f"RXCHFmult_GAM2_IC1ex {what}")
    for which in range(3):
        n = npb if which == 1 else ne
        F = np.zeros((n, n), order="F")
        e = oracle.rxchfmult_omgex(2, which, ne, npb, DAE, DBE, DP, x2[1], x2[2], None, F)
        check_close(parts[87 + 2 * which], F.ravel(order="F"), what=f"RXCHFmult omg2ex {which}")
        assert parts[88 + 2 * which][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    x3 = oracle.rxchfmult_gam3_ic1ex(s)
    for k, what in enumerate(("INT_GAM3", "INT_GAM3ex1", "INT_GAM3ex2")):
        check_close(parts[93 + k], x3[k], what=f"RXCHFmult_GAM3_IC1ex {what}")
    for which in range(3):
        n = npb if which == 1 else ne
        F = np.zeros((n, n), order="F")
        e = oracle.rxchfmult_omgex(3, which, ne, npb, DAE, DBE, DP, x3[0], x3[1], x3[2], F)
        check_close(parts[96 + 2 * which], F.ravel(order="F"), what=f"RXCHFmult omg3ex {which}")
        assert parts[97 + 2 * which][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    # RXCHF with one special electron: DAE regular, DBE special
    r1 = oracle.rxchf_gam2_ic1(s)
    for k, what in enumerate(("GM2_1ICR", "GM2_2ICR", "GM2sICR")):
        check_close(parts[102 + k], r1[k], what=f"RXCHF_GAM2_IC1 {what}")
    for which in range(3):
        n = npb if which == 1 else ne
        F, SF = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        e, sv = oracle.rxchf_omg2(which, ne, npb, DAE, DBE, DP, *r1, F, SF)
        at = 105 + 4 * which
        check_close(parts[at], F.ravel(order="F"), what=f"RXCHF omg2 {which} F")
        check_close(parts[at + 1], SF.ravel(order="F"), what=f"RXCHF omg2 {which} S")
        assert parts[at + 2][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
        assert parts[at + 3][0] == pytest.approx(sv, rel=1e-11, abs=1e-14)
    # RXCHF e-e-e-p and GAM_1 pieces
    r3 = oracle.rxchf_gam3_ic1(s)
    check_close(parts[117], r3[0], what="RXCHF_GAM3_IC1 GM3_1ICR")
    check_close(parts[118], r3[1], what="RXCHF_GAM3_IC1 GM3_2ICR")
    for which in range(3):
        n = npb if which == 1 else ne
        F = np.zeros((n, n), order="F")
        e = oracle.rxchf_omg3(which, ne, npb, DAE, DBE, DP, *r3, F)
        check_close(parts[119 + 2 * which], F.ravel(order="F"), what=f"RXCHF omg3 {which}")
        assert parts[120 + 2 * which][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    rg1, rg1s = oracle.rxchfmult_gam1(s)
    for k, which in enumerate((1, 2)):
        n = npb if which == 1 else ne
        F, SF = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        e, sv = oracle.rxchf_omg1(which, ne, npb, DBE, DP, rg1, rg1s, F, SF)
        at = 125 + 4 * k
        check_close(parts[at], F.ravel(order="F"), what=f"RXCHF omg1 {which} F")
        check_close(parts[at + 1], SF.ravel(order="F"), what=f"RXCHF omg1 {which} S")
        assert parts[at + 2][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
        assert parts[at + 3][0] == pytest.approx(sv, rel=1e-11, abs=1e-14)
    # RXCUHF over the RXCHF tensors: alpha DAE, beta DBE, total DE, special DBE
    r2 = oracle.rxchf_gam2_ic1(s)
    F, F2, SF = (np.zeros((ne, ne), order="F") for _ in range(3))
    e, sv = oracle.rxcuhf_omg2(0, ne, npb, DAE, DE, DBE, DP, *r2, F, SF, F2)
    check_close(parts[133], F.ravel(order="F"), what="RXCUHF_FE_OMG2 FAalpE")
    check_close(parts[134], F2.ravel(order="F"), what="RXCUHF_FE_OMG2 FAbetE")
    check_close(parts[135], SF.ravel(order="F"), what="RXCUHF_FE_OMG2 SAalpE")
    assert parts[136][0] == pytest.approx(e, rel=1e-11, abs=1e-14) and parts[137][0] == pytest.approx(sv, rel=1e-11, abs=1e-14)
    for k, which in enumerate((1, 2)):
        n = npb if which == 1 else ne
        F, SF = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        e, sv = oracle.rxcuhf_omg2(which, ne, npb, DAE, DE, DBE, DP, *r2, F, SF)
        at = 138 + 4 * k
        check_close(parts[at], F.ravel(order="F"), what=f"RXCUHF omg2 {which} F")
        check_close(parts[at + 1], SF.ravel(order="F"), what=f"RXCUHF omg2 {which} S")
        assert parts[at + 2][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
        assert parts[at + 3][0] == pytest.approx(sv, rel=1e-11, abs=1e-14)
    F, F2 = np.zeros((ne, ne), order="F"), np.zeros((ne, ne), order="F")
    e = oracle.rxcuhf_omg3(0, ne, npb, DAE, DBE, DE, DBE, DP, *r3, F, F2)
    check_close(parts[146], F.ravel(order="F"), what="RXCUHF_FE_OMG3 FAalpE")
    check_close(parts[147], F2.ravel(order="F"), what="RXCUHF_FE_OMG3 FAbetE")
    assert parts[148][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
    for k, which in enumerate((1, 2)):
        n = npb if which == 1 else ne
        F = np.zeros((n, n), order="F")
        e = oracle.rxcuhf_omg3(which, ne, npb, DAE, DBE, DE, DBE, DP, *r3, F)
        check_close(parts[149 + 2 * k], F.ravel(order="F"), what=f"RXCUHF omg3 {which}")
        assert parts[150 + 2 * k][0] == pytest.approx(e, rel=1e-11, abs=1e-14)
