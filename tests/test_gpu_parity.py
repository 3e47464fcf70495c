"""GPU parity: libneorx_hip.so through its C ABI against the oracle on the
same seeded inputs (FP64 bar in tests/_util.py)."""
import numpy as np
import pytest

import neorx
import oracle
from neorx import synth
from _util import check_close, random_tuples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    torch.cuda.init()
    arch = neorx.init(0)
    assert arch.startswith("gfx950"), arch
    return torch


def _run_list(torch, S, cls, idx):
    n = len(idx)
    d_idx = torch.from_numpy(np.ascontiguousarray(idx)).cuda()
    d_e = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_s = torch.zeros(n, dtype=torch.float64, device="cuda")
    S.prim_list_dev(cls, n, d_idx.data_ptr(), d_e.data_ptr(), d_s.data_ptr(),
                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return d_e.cpu().numpy(), d_s.cpu().numpy()


@pytest.mark.parametrize("cls,n", [(1, 2048), (2, 1024), (3, 256)])
def test_prim_list_vs_oracle(torch_dev, cls, n):
    print(f"case prim_list cls={cls} n={n}", flush=True)
    s = synth.small(seed=60 + cls, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    S = neorx.System(s)
    idx = random_tuples(s, cls, n, seed=cls)
    ge, gs = _run_list(torch_dev, S, cls, idx)
    oe, os_ = oracle.prim_list(s, cls, idx)
    st = check_close(ge, oe, what=f"cls{cls} E")
    print(f"  cls{cls} E: {st}")
    if cls < 3:
        check_close(gs, os_, what=f"cls{cls} S")
    S.close()


def test_prim_config2_sample(torch_dev):
    s = synth.config2()
    S = neorx.System(s)
    for cls, n in [(1, 1024), (2, 512), (3, 96)]:
        print(f"case config2 sample cls={cls} n={n}", flush=True)
        idx = random_tuples(s, cls, n, seed=100 + cls)
        ge, gs = _run_list(torch_dev, S, cls, idx)
        oe, os_ = oracle.prim_list(s, cls, idx)
        check_close(ge, oe, what=f"config2 cls{cls} E")
    S.close()


def test_prim_range_matches_enumeration(torch_dev):
    torch = torch_dev
    s = synth.small(seed=70, e_shells=(0, 1, 2), p_shells=(0, 1))
    S = neorx.System(s)
    eo, po = S.order()
    ne, npb = s["npebf"], s["npbf"]
    start, count = 12345, 3000
    d_e = torch.zeros(count, dtype=torch.float64, device="cuda")
    d_s = torch.zeros(count, dtype=torch.float64, device="cuda")
    S.prim_range_dev(2, start, count, d_e.data_ptr(), d_s.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    r = np.arange(start, start + count)
    je2 = r % ne; r //= ne
    ie2 = r % ne; r //= ne
    je1 = r % ne; r //= ne
    ie1 = r % ne; r //= ne
    jp = r % npb; ip = r // npb
    idx = np.stack([eo[ie1], eo[ie2], po[ip], eo[je1], eo[je2], po[jp]], axis=1)
    oe, os_ = oracle.prim_list(s, 2, idx)
    check_close(d_e.cpu().numpy(), oe, what="range E")
    check_close(d_s.cpu().numpy(), os_, what="range S")
    S.close()


def test_gam1_gam2_builders(torch_dev):
    s = synth.small(seed=71, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1))
    S = neorx.System(s)
    g1, g1s = S.GAM1()
    o1, o1s = oracle.gam1(s)
    check_close(g1, o1, what="GAM_1")
    check_close(g1s, o1s, what="GAM_1s")
    g2, g2s = S.GAM2_IC1()
    o2, o2s = oracle.gam2_ic1(s)
    check_close(g2, o2, what="GM2ICR")
    check_close(g2s, o2s, what="GM2sICR")
    S.close()


def test_resident_tensors(torch_dev):
    """The host-pointer builders keep GM2ICR/GM2sICR/GM3ICR on the device: the
    host-pointer contractions given the same arrays copy no tensor over PCIe
    and agree with the oracle; changing the caller's array drops the copy (the
    next call copies the new values and agrees with the oracle on them)."""
    s = synth.small(seed=77, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1))
    ne, npb = s["nebf"], s["npbf"]
    print(f"case resident tensors nebf={ne} npbf={npb}", flush=True)
    S = neorx.System(s)
    g2, g2s = S.GAM2_IC1()
    g3 = S.GAM3_IC1()
    S.close()
    DE, DP = synth.density(ne, 4.0, 3), synth.density(npb, 1.0, 4)
    F0e = np.asfortranarray(np.random.default_rng(1).standard_normal((ne, ne)))
    F0p = np.asfortranarray(np.random.default_rng(2).standard_normal((npb, npb)))

    def both(g, fn):
        A, B = F0e.copy(order="F"), F0e.copy(order="F")
        P, Q = F0p.copy(order="F"), F0p.copy(order="F")
        h0 = neorx.transfer_bytes()[0]
        eg = fn(neorx, A, P)
        dh = neorx.transfer_bytes()[0] - h0
        eo = fn(oracle, B, Q)
        check_close(A, B, what="FE")
        check_close(P, Q, what="FP")
        return eg, eo, dh

    eg, eo, dh = both(g3, lambda m, A, P: m.eg3ic1(ne, npb, DE, DP, g3, A, P) if m is oracle
                      else m.EG3IC1(ne, npb, DE, DP, g3, A, P))
    assert eg == pytest.approx(eo, rel=1e-11) and dh == 0, (eg, eo, dh)
    r = both(g2, lambda m, A, P: m.eg2ic1(ne, npb, DE, DP, g2, g2s, A, P)[2] if m is oracle
             else m.EG2IC1(ne, npb, DE, DP, g2, g2s, A, P)[2])
    assert r[0] == pytest.approx(r[1], rel=1e-11) and r[2] == 0, r
    g3[0] += 0.5  # a probed entry
    eg, eo, dh = both(g3, lambda m, A, P: m.eg3ic1(ne, npb, DE, DP, g3, A, P) if m is oracle
                      else m.EG3IC1(ne, npb, DE, DP, g3, A, P))
    assert eg == pytest.approx(eo, rel=1e-11) and dh == g3.size * 8, (eg, eo, dh)
    neorx.release_resident()


@pytest.mark.parametrize("kw", [
    dict(seed=72, n_e_centres=3, e_shells=(0,), contract=(1,), p_shells=(0, 1)),
    dict(seed=74, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,)),
    dict(seed=75, n_e_centres=1, e_shells=(2,), contract=(1,), p_shells=(1,)),
], ids=["s-shells", "sp-contracted", "d-electrons-p-nucleus"])
def test_gam3_builder(torch_dev, kw):
    """GAM3_IC1: primitive tuples, contraction over primitives of length 1-2
    and the six-permutation average, s/p/d electrons."""
    s = synth.small(**kw)
    print(f"case gam3 builder nebf={s['nebf']} npebf={s['npebf']} npbf={s['npbf']}", flush=True)
    S = neorx.System(s)
    g3 = S.GAM3_IC1()
    o3 = oracle.gam3_ic1(s)
    check_close(g3, o3, what="GM3ICR")
    S.close()


def test_builders_grid_stride(torch_dev):
    """The full-tensor symmetrisation kernels stride over their entries when
    the grid is capped: GAM2_IC1 and GAM3_IC1 with a one-block grid still
    equal the oracle."""
    import os

    s = synth.small(seed=76, n_e_centres=1, e_shells=(0, 1), contract=(1,), p_shells=(0, 1))
    ne, npb = s["nebf"], s["npbf"]
    print(f"case grid cap nebf={ne} npbf={npb}", flush=True)
    # one 256-lane block: GM2 (ne^4 npbf^2 entries) and each (ip,jp) block of
    # GM3 (ne^6 entries) take several passes of the stride loop
    assert ne ** 4 * npb ** 2 >= 8 * 256 and ne ** 6 >= 8 * 256
    os.environ["NEORX_GRID_CAP"] = "1"
    try:
        neorx.init(0)
        S = neorx.System(s)
        g2, g2s = S.GAM2_IC1()
        g3 = S.GAM3_IC1()
        S.close()
    finally:
        del os.environ["NEORX_GRID_CAP"]
        neorx.init(0)
    o2, o2s = oracle.gam2_ic1(s)
    check_close(g2, o2, what="GM2ICR (capped grid)")
    check_close(g2s, o2s, what="GM2sICR (capped grid)")
    check_close(g3, oracle.gam3_ic1(s), what="GM3ICR (capped grid)")


def test_gam_ee_and_gam4(torch_dev):
    s = synth.small(seed=73, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,))
    S = neorx.System(s)
    gee = S.GAM_ee()
    oee = oracle.gam_ee(s)
    check_close(gee, oee, what="GAM_ee")
    _, g2s = oracle.gam2_ic1(s)
    ne, npb = s["nebf"], s["npbf"]
    sub = 3  # in-core GAM_4 grows as nebf^8: use the first 3 functions
    g2s_sub = g2s.reshape(npb, npb, ne, ne, ne, ne)[:, :, :sub, :sub, :sub, :sub].copy()
    gee_sub = oee.reshape(ne, ne, ne, ne)[:sub, :sub, :sub, :sub].copy()
    g4 = neorx.GAM4_ICR(sub, npb, g2s_sub, gee_sub)
    o4 = oracle.gam4_icr(sub, npb, g2s_sub, gee_sub)
    check_close(g4, o4, what="GM4ICR")
    S.close()


@pytest.mark.parametrize("ne,npb", [(5, 3), (24, 8)])
def test_eg2ic1(torch_dev, ne, npb):
    rng = np.random.default_rng(ne)
    G = rng.standard_normal(ne ** 4 * npb * npb)
    Gs = rng.standard_normal(ne ** 4 * npb * npb)
    DE, DP = synth.density(ne, 14.0, 1), synth.density(npb, 1.0, 2)
    FE0 = np.asfortranarray(rng.standard_normal((ne, ne)))
    FP0 = np.asfortranarray(rng.standard_normal((npb, npb)))
    FE1, FP1, FE2, FP2 = (x.copy(order="F") for x in (FE0, FP0, FE0, FP0))
    a = neorx.EG2IC1(ne, npb, DE, DP, G, Gs, FE1, FP1)
    b = oracle.eg2ic1(ne, npb, DE, DP, G, Gs, FE2, FP2)
    check_close(FE1, FE2, what="FE")
    check_close(FP1, FP2, what="FP")
    check_close(a[0], b[0], what="SE")
    check_close(a[1], b[1], what="SP")
    assert a[2] == pytest.approx(b[2], rel=1e-11)
    assert a[3] == pytest.approx(b[3], rel=1e-11)


def test_eg3_eg4(torch_dev):
    rng = np.random.default_rng(9)
    ne, npb = 4, 2
    G3 = rng.standard_normal(ne ** 6 * npb * npb)
    DE, DP = synth.density(ne, 3.0, 3), synth.density(npb, 1.0, 4)
    FE0 = np.asfortranarray(rng.standard_normal((ne, ne)))
    FP0 = np.asfortranarray(rng.standard_normal((npb, npb)))
    F = [x.copy(order="F") for x in (FE0, FP0, FE0, FP0)]
    e1 = neorx.EG3IC1(ne, npb, DE, DP, G3, F[0], F[1])
    e2 = oracle.eg3ic1(ne, npb, DE, DP, G3, F[2], F[3])
    check_close(F[0], F[2], what="EG3 FE")
    check_close(F[1], F[3], what="EG3 FP")
    assert e1 == pytest.approx(e2, rel=1e-11)
    ne = 3
    G4 = rng.standard_normal(ne ** 8 * npb * npb)
    DE = synth.density(ne, 3.0, 5)
    F = [x.copy(order="F") for x in (FE0[:ne, :ne], FP0, FE0[:ne, :ne], FP0)]
    e1 = neorx.EG4ICR(ne, npb, DE, DP, G4, F[0], F[1])
    e2 = oracle.eg4icr(ne, npb, DE, DP, G4, F[2], F[3])
    check_close(F[0], F[2], what="EG4 FE")
    check_close(F[1], F[3], what="EG4 FP")
    assert e1 == pytest.approx(e2, rel=1e-11)


@pytest.mark.parametrize("ne", [33, 46])
def test_eg2_eg4_fixed_order(torch_dev, ne):
    """EG2IC1 at basis sizes that take the wider launch shapes of its slab
    kernel (nebf 33: 256 lanes x 8 columns; 46: 1024 lanes), and EG4ICR:
    against the oracle, and two runs of the same call bitwise equal (every
    partial sum has a fixed order, no floating-point atomics)."""
    print(f"case eg2/eg4 fixed order ne={ne}", flush=True)
    rng = np.random.default_rng(ne)
    npb = 1
    G, Gs = rng.standard_normal(ne ** 4), rng.standard_normal(ne ** 4)
    DE, DP = synth.density(ne, 14.0, 1), synth.density(npb, 1.0, 2)
    FE0 = np.asfortranarray(rng.standard_normal((ne, ne)))
    runs = []
    for _ in range(2):
        FE, FP = FE0.copy(order="F"), np.zeros((npb, npb), order="F")
        r = neorx.EG2IC1(ne, npb, DE, DP, G, Gs, FE, FP)
        runs.append([FE, FP, r[0], r[1], np.array(r[2:4])])
    for a, b in zip(*runs):
        assert np.array_equal(a, b), "EG2IC1: two runs differ"
    FE, FP = FE0.copy(order="F"), np.zeros((npb, npb), order="F")
    o = oracle.eg2ic1(ne, npb, DE, DP, G, Gs, FE, FP)
    check_close(runs[0][0], FE, what="FE")
    check_close(runs[0][1], FP, what="FP")
    check_close(runs[0][2], o[0], what="SE")
    assert runs[0][4][0] == pytest.approx(o[2], rel=1e-11)
    ne4, npb4 = 4, 2
    G4 = rng.standard_normal(ne4 ** 8 * npb4 * npb4)
    DE4, DP4 = synth.density(ne4, 3.0, 5), synth.density(npb4, 1.0, 4)
    runs = []
    for _ in range(2):
        FE, FP = np.zeros((ne4, ne4), order="F"), np.zeros((npb4, npb4), order="F")
        e = neorx.EG4ICR(ne4, npb4, DE4, DP4, G4, FE, FP)
        runs.append([FE, FP, np.array([e])])
    for a, b in zip(*runs):
        assert np.array_equal(a, b), "EG4ICR: two runs differ"
    FE, FP = np.zeros((ne4, ne4), order="F"), np.zeros((npb4, npb4), order="F")
    e = oracle.eg4icr(ne4, npb4, DE4, DP4, G4, FE, FP)
    check_close(runs[0][0], FE, what="EG4 FE")
    check_close(runs[0][1], FP, what="EG4 FP")
    assert runs[0][2][0] == pytest.approx(e, rel=1e-11)


def test_direct_scf_routines(torch_dev):
    """GAM2/3/4_DSCF through the C ABI against the oracle's restatement of the
    reference routines; Fock matrices start non-zero to check accumulation."""
    s = synth.small(seed=71, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    ne, npb = s["nebf"], s["npbf"]
    print(f"case direct SCF ne={ne} np={npb}", flush=True)
    DE, DP = synth.density(ne, 2.0, 3), synth.density(npb, 1.0, 4)
    rng = np.random.default_rng(5)
    F0 = [np.asfortranarray(rng.standard_normal((n, n))) for n in (ne, npb, ne, npb)]
    S = neorx.System(s)
    G = [f.copy(order="F") for f in F0]
    O = [f.copy(order="F") for f in F0]
    ge, gs = S.GAM2_DSCF(DE, DP, *G)
    oe, os_ = oracle.gam2_dscf(s, DE, DP, *O)
    for g, o, nm in zip(G, O, ("FE", "FP", "SE2", "SP2")):
        check_close(g, o, what=f"GAM2_DSCF {nm}")
    assert ge == pytest.approx(oe, rel=1e-11) and gs == pytest.approx(os_, rel=1e-11)
    G = [f.copy(order="F") for f in F0[:2]]
    O = [f.copy(order="F") for f in F0[:2]]
    ge3 = S.GAM3_DSCF(DE, DP, *G)
    oe3 = oracle.gam3_dscf(s, DE, DP, *O)
    for g, o, nm in zip(G, O, ("FE", "FP")):
        check_close(g, o, what=f"GAM3_DSCF {nm}")
    assert ge3 == pytest.approx(oe3, rel=1e-11)
    og2, og2s = oracle.gam2_ic1(s)
    oee = oracle.gam_ee(s)
    G = [f.copy(order="F") for f in F0[:2]]
    O = [f.copy(order="F") for f in F0[:2]]
    ge4 = neorx.GAM4_DSCF(ne, npb, og2s, oee, DE, DP, *G)
    oe4 = oracle.gam4_dscf(ne, npb, og2s, oee, DE, DP, *O)
    for g, o, nm in zip(G, O, ("FE", "FP")):
        check_close(g, o, what=f"GAM4_DSCF {nm}")
    assert ge4 == pytest.approx(oe4, rel=1e-11)
    S.close()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("ne,npb", [(5, 3), (12, 2)])
def test_mixed_precision_contractions(torch_dev, ne, npb):
    """FP32-stored tensors with FP64 arithmetic: the GPU result equals the
    oracle's FP64 contraction of the FP32-rounded tensor (bar of tests/_util),
    and stays within 1e-6 (relative, per matrix) of the FP64 path."""
    torch = torch_dev
    print(f"case mixed precision ne={ne} np={npb}", flush=True)
    rng = np.random.default_rng(ne)
    DE, DP = synth.density(ne, 2.0, 7), synth.density(npb, 1.0, 8)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    g3r = g3.astype(np.float32).astype(np.float64)
    d_de, d_dp = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F"))
    d_g3 = _dev(torch, g3)
    d_g3f = torch.empty(g3.size, dtype=torch.float32, device="cuda")
    neorx.call_dev("neorx_convert_f32_dev", g3.size, d_g3.data_ptr(), d_g3f.data_ptr(), 0)
    res = {}
    for name, gp in (("neorx_eg3ic1_dev", d_g3), ("neorx_eg3ic1_f32_dev", d_g3f)):
        fe = torch.zeros(ne * ne, dtype=torch.float64, device="cuda")
        fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        neorx.call_dev(name, ne, npb, d_de.data_ptr(), d_dp.data_ptr(), gp.data_ptr(), fe.data_ptr(), fp.data_ptr(),
                       e.data_ptr(), 0)
        torch.cuda.synchronize()
        res[name] = (fe.cpu().numpy(), fp.cpu().numpy(), float(e.cpu()[0]))
    for name, g in (("neorx_eg3ic1_dev", g3), ("neorx_eg3ic1_f32_dev", g3r)):
        FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
        E = oracle.eg3ic1(ne, npb, DE, DP, g, FE, FP)
        check_close(res[name][0], FE.ravel(order="F"), what=f"{name} FE")
        check_close(res[name][1], FP.ravel(order="F"), what=f"{name} FP")
        assert res[name][2] == pytest.approx(E, rel=1e-11)
    a, b = res["neorx_eg3ic1_dev"], res["neorx_eg3ic1_f32_dev"]
    assert np.max(np.abs(a[0] - b[0])) <= 1e-6 * np.max(np.abs(a[0]))
    if ne <= 5:
        g2, g2s = rng.standard_normal(ne ** 4 * npb * npb), rng.standard_normal(ne ** 4 * npb * npb)
        g4 = rng.standard_normal(ne ** 8 * npb * npb)
        d = {k: _dev(torch, v) for k, v in (("g2", g2), ("g2s", g2s), ("g4", g4))}
        f = {k: torch.empty(v.numel(), dtype=torch.float32, device="cuda") for k, v in d.items()}
        for k in d:
            neorx.call_dev("neorx_convert_f32_dev", d[k].numel(), d[k].data_ptr(), f[k].data_ptr(), 0)
        rd = lambda x: x.astype(np.float32).astype(np.float64)  # noqa: E731
        out = [torch.zeros(n, dtype=torch.float64, device="cuda") for n in (ne * ne, ne * ne, npb * npb, npb * npb, 2)]
        neorx.call_dev("neorx_eg2ic1_f32_dev", ne, npb, d_de.data_ptr(), d_dp.data_ptr(), f["g2"].data_ptr(),
                       f["g2s"].data_ptr(), out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                       out[3].data_ptr(), out[4].data_ptr(), out[4].data_ptr() + 8, 0)
        torch.cuda.synchronize()
        FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
        SE, SP, E, S = oracle.eg2ic1(ne, npb, DE, DP, rd(g2), rd(g2s), FE, FP)
        check_close(out[0].cpu().numpy(), FE.ravel(order="F"), what="EG2 f32 FE")
        check_close(out[1].cpu().numpy(), SE.ravel(order="F"), what="EG2 f32 SE")
        check_close(out[2].cpu().numpy(), FP.ravel(order="F"), what="EG2 f32 FP")
        assert out[4].cpu().numpy()[0] == pytest.approx(E, rel=1e-11)
        fe = torch.zeros(ne * ne, dtype=torch.float64, device="cuda")
        fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        neorx.call_dev("neorx_eg4icr_f32_dev", ne, npb, d_de.data_ptr(), d_dp.data_ptr(), f["g4"].data_ptr(),
                       fe.data_ptr(), fp.data_ptr(), e.data_ptr(), 0)
        torch.cuda.synchronize()
        FE, FP = np.zeros((ne, ne), order="F"), np.zeros((npb, npb), order="F")
        E4 = oracle.eg4icr(ne, npb, DE, DP, rd(g4), FE, FP)
        check_close(fe.cpu().numpy(), FE.ravel(order="F"), what="EG4 f32 FE")
        check_close(fp.cpu().numpy(), FP.ravel(order="F"), what="EG4 f32 FP")
        assert float(e.cpu()[0]) == pytest.approx(E4, rel=1e-11)


@pytest.mark.parametrize("ne,npb", [(16, 1), (24, 1), (25, 1), (16, 2)])
def test_eg3_wide_slabs(torch_dev, ne, npb):
    """EG3IC1 at basis sizes whose slab loop runs the 16-wide (FP64) and
    12-wide (FP32) load blocks, the 8-wide block and the scalar remainder:
    nebf 16 = 16 (FP64) / 12+4 (FP32), 24 = 16+8 / 12+12, 25 = 16+8+1 /
    12+12+1.  FP64 and FP32 storage against the oracle, and two runs of the
    same call bitwise equal (fixed-order reduction)."""
    torch = torch_dev
    print(f"case eg3 wide ne={ne} np={npb}", flush=True)
    rng = np.random.default_rng(100 + ne + npb)
    DE, DP = synth.density(ne, 10.0, ne), synth.density(npb, 1.0, 3)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    d_de, d_dp = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F"))
    d_g3 = _dev(torch, g3)
    d_g3f = torch.empty(g3.size, dtype=torch.float32, device="cuda")
    neorx.call_dev("neorx_convert_f32_dev", g3.size, d_g3.data_ptr(), d_g3f.data_ptr(), 0)
    torch.cuda.synchronize()
    FE0 = rng.standard_normal((ne, ne))
    for name, gp, gh in (("neorx_eg3ic1_dev", d_g3, g3),
                         ("neorx_eg3ic1_f32_dev", d_g3f, g3.astype(np.float32).astype(np.float64))):
        outs = []
        for rep in range(2):
            fe = _dev(torch, FE0.ravel(order="F"))
            fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
            e = torch.zeros(1, dtype=torch.float64, device="cuda")
            neorx.call_dev(name, ne, npb, d_de.data_ptr(), d_dp.data_ptr(), gp.data_ptr(), fe.data_ptr(),
                           fp.data_ptr(), e.data_ptr(), 0)
            torch.cuda.synchronize()
            outs.append((fe.cpu().numpy(), fp.cpu().numpy(), e.cpu().numpy()))
        for a, b in zip(outs[0], outs[1]):
            assert np.array_equal(a, b), f"{name}: two runs differ"
        FE, FP = np.asfortranarray(FE0.copy()), np.zeros((npb, npb), order="F")
        E = oracle.eg3ic1(ne, npb, DE, DP, gh, FE, FP)
        check_close(outs[0][0], FE.ravel(order="F"), what=f"{name} FE")
        check_close(outs[0][1], FP.ravel(order="F"), what=f"{name} FP")
        assert outs[0][2][0] == pytest.approx(E, rel=1e-11)


def test_autotune_eg3(torch_dev):
    """neorx_autotune_eg3_dev times the 8, 12, 16 and 24 loads-in-flight
    variants of the EG3 slab kernel for each storage type and records the
    fastest; the slab sums run over c in the same order whatever the depth,
    so the tuned launch gives bitwise the results of the default one."""
    torch = torch_dev
    ne, npb = 16, 2
    print(f"case autotune eg3 ne={ne} np={npb}", flush=True)
    rng = np.random.default_rng(77)
    DE, DP = synth.density(ne, 6.0, 7), synth.density(npb, 1.0, 8)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    d_de, d_dp, d_g3 = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F")), _dev(torch, g3)
    d_g3f = torch.empty(g3.size, dtype=torch.float32, device="cuda")
    neorx.call_dev("neorx_convert_f32_dev", g3.size, d_g3.data_ptr(), d_g3f.data_ptr(), 0)
    torch.cuda.synchronize()

    def run(name, g):
        fe = torch.zeros(ne * ne, dtype=torch.float64, device="cuda")
        fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
        e = torch.zeros(1, dtype=torch.float64, device="cuda")
        neorx.call_dev(name, ne, npb, d_de.data_ptr(), d_dp.data_ptr(), g.data_ptr(), fe.data_ptr(), fp.data_ptr(),
                       e.data_ptr(), 0)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (fe, fp, e)]

    for dtype, name, g in ((0, "neorx_eg3ic1_dev", d_g3), (1, "neorx_eg3ic1_f32_dev", d_g3f)):
        before = run(name, g)
        best, ms, allms = neorx.autotune_eg3_dev(ne, npb, dtype, d_de.data_ptr(), g.data_ptr(), reps=2)
        print(f"  dtype {dtype}: best {best} ({ms:.3f} ms), all {allms}", flush=True)
        assert best in (8, 12, 16, 24) and ms > 0 and all(m > 0 for m in allms)
        assert ms == pytest.approx(min(allms), rel=1e-6)
        after = run(name, g)
        for a, b in zip(before, after):
            assert np.array_equal(a, b)
    d_g3h = torch.from_numpy(g3.astype(np.float32)).cuda().to(torch.bfloat16)
    best, ms, allms = neorx.autotune_eg3_dev(ne, npb, 2, d_de.data_ptr(), d_g3h.data_ptr(), reps=2)
    assert best in (8, 12, 16, 24) and ms > 0
    with pytest.raises(neorx.NeorxError):
        neorx.autotune_eg3_dev(ne, npb, 3, d_de.data_ptr(), d_g3.data_ptr(), reps=1)


def test_precision_switch(torch_dev):
    """FP64 is selected once successive densities differ by less than the
    threshold, and stays selected."""
    torch = torch_dev
    ne = 7
    rng = np.random.default_rng(3)
    D0 = synth.density(ne, 2.0, 9)
    state = torch.zeros(ne * ne + 2, dtype=torch.float64, device="cuda")
    seen = []
    for k in range(1, 10):
        noise = rng.uniform(-1, 1, (ne, ne))
        D = D0 + 10.0 ** (-k) * (noise + noise.T)
        d = _dev(torch, D.ravel(order="F"))
        use, delta = neorx.precision_switch_dev(ne, d.data_ptr(), state.data_ptr(), 1e-5)
        if k == 1:
            assert delta == float("inf") and not use
        else:
            ref = np.max(np.abs(D - prev))
            assert delta == pytest.approx(ref, rel=1e-15)
        seen.append(use)
        prev = D
    first = seen.index(True)
    assert all(seen[first:]) and not any(seen[:first])
    assert 4 <= first <= 7


@pytest.mark.parametrize("nae,nbe", [(1, 1), (3, 2), (4, 4), (2, 3)])
def test_uhf_contractions(torch_dev, nae, nbe):
    """FAE_OMG2/3/4, UFP_OMG2/3/4 through the C ABI against the oracle's
    restatement of source/xcuhf/xcuhf_omg{2,3,4}.f, electron counts chosen to
    switch the spin-pair terms on and off; Fock matrices start non-zero."""
    print(f"case uhf nae={nae} nbe={nbe}", flush=True)
    rng = np.random.default_rng(20 + 10 * nae + nbe)
    ne, npb = 4, 2
    DAE, DBE = synth.density(ne, float(nae), 5), synth.density(ne, float(nbe), 6)
    DP = synth.density(npb, 1.0, 7)
    g2, g2s = rng.standard_normal(ne ** 4 * npb * npb), rng.standard_normal(ne ** 4 * npb * npb)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    g4 = rng.standard_normal(ne ** 8 * npb * npb)
    a = (nae, nbe, ne, npb, DAE, DBE, DP)
    for gfn, ofn, n in ((neorx.FAE_OMG2, oracle.fae_omg2, ne), (neorx.UFP_OMG2, oracle.ufp_omg2, npb)):
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        Fg, Fo, So = F0.copy(order="F"), F0.copy(order="F"), np.zeros((n, n), order="F")
        Sg, eg, sg = gfn(*a, g2, g2s, Fg)
        eo, so = ofn(*a, g2, g2s, Fo, So)
        check_close(Fg, Fo, what=gfn.__name__)
        check_close(Sg, So, what=gfn.__name__ + " S")
        assert eg == pytest.approx(eo, rel=1e-11, abs=1e-14) and sg == pytest.approx(so, rel=1e-11, abs=1e-14)
    for gfn, ofn, G, n in ((neorx.FAE_OMG3, oracle.fae_omg3, g3, ne), (neorx.UFP_OMG3, oracle.ufp_omg3, g3, npb),
                           (neorx.FAE_OMG4, oracle.fae_omg4, g4, ne), (neorx.UFP_OMG4, oracle.ufp_omg4, g4, npb)):
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        Fg, Fo = F0.copy(order="F"), F0.copy(order="F")
        eg = gfn(*a, G, Fg)
        eo = ofn(*a, G, Fo)
        check_close(Fg, Fo, what=gfn.__name__)
        assert eg == pytest.approx(eo, rel=1e-11, abs=1e-14), gfn.__name__


def test_omg4_icr(torch_dev):
    s = synth.small(seed=16, n_e_centres=2, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    g2, g2s = oracle.gam2_ic1(s)
    gee = oracle.gam_ee(s)
    check_close(neorx.OMG4_ICR(s["nebf"], s["npbf"], g2s, gee), oracle.omg4_icr(s["nebf"], s["npbf"], g2s, gee),
                what="OMG4_ICR")


def test_rxchfne(torch_dev):
    """RXCHF-ne: class-4 tuples, RXCHFne_GAM2_IC1 and the RXCHF-ne /
    RXCUHF-ne contractions through the C ABI against the oracle."""
    torch = torch_dev
    s = synth.small(seed=48, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    S = neorx.System(s)
    idx = random_tuples(s, 2, 512, seed=9)
    ge, gs = _run_list(torch, S, 4, idx)
    oe, os_ = oracle.prim_list(s, 4, idx)
    check_close(ge, oe, what="RXCHF-ne ansE1")
    check_close(gs, os_, what="RXCHF-ne ansE2")
    S.close()
    t = synth.small(seed=49, n_e_centres=2, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    T = neorx.System(t)
    g, gx = T.RXCHFne_GAM2_IC1()
    og, ogx = oracle.rxchfne_gam2_ic1(t)
    check_close(g, og, what="RXCHFne_GAM2_IC1 GM2")
    check_close(gx, ogx, what="RXCHFne_GAM2_IC1 GM2ex")
    T.close()
    ne, npb = t["nebf"], t["npbf"]
    rng = np.random.default_rng(12)
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    DAa = synth.density(ne, 1.0, 4)
    for k, which in enumerate(("fae", "fp", "fbe")):
        n = npb if which == "fp" else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        eg = neorx.RXCHFne_OMG2(which, ne, npb, DA, DB, DP, og, A)
        eo = oracle.rxchfne_omg2(k, ne, npb, DA, DB, DP, og, B)
        check_close(A, B, what=f"RXCHFne_{which}_OMG2")
        assert eg == pytest.approx(eo, rel=1e-11, abs=1e-15)
        wu = {"fae": "fe", "fp": "fp", "fbe": "fbe"}[which]
        A, B = F0.copy(order="F"), F0.copy(order="F")
        F20 = np.asfortranarray(rng.standard_normal((ne, ne)))
        A2, B2 = F20.copy(order="F"), F20.copy(order="F")
        eg = neorx.RXCUHFne_OMG2(wu, ne, npb, DAa, DA, DB, DP, og, ogx, A, A2)
        eo = oracle.rxcuhfne_omg2(k, ne, npb, DAa, DA, DB, DP, og, ogx, B, B2)
        check_close(A, B, what=f"RXCUHFne_{wu}_OMG2")
        if wu == "fe":
            check_close(A2, B2, what="RXCUHFne_FE_OMG2 beta")
        assert eg == pytest.approx(eo, rel=1e-11, abs=1e-15)
        # fixed-order sums: a second call agrees bit for bit
        A3, A4 = F0.copy(order="F"), F20.copy(order="F")
        assert neorx.RXCUHFne_OMG2(wu, ne, npb, DAa, DA, DB, DP, og, ogx, A3, A4) == eg
        assert np.array_equal(A3, A) and np.array_equal(A4, A2)
    gee = oracle.gam_ee(t)
    F0 = np.asfortranarray(rng.standard_normal((ne, ne)))
    A, B = F0.copy(order="F"), F0.copy(order="F")
    assert neorx.RXCHFne_FAE_GAMee(ne, gee, DA, A) == pytest.approx(oracle.rxchfne_fae_gamee(ne, gee, DA, B),
                                                                   rel=1e-11)
    check_close(A, B, what="RXCHFne_FAE_GAMee")


def _round_storage(g, prec):
    """A tensor as the device stores it in FP32 or BF16 (round to nearest even
    through FP32), widened back to FP64."""
    f = np.asarray(g, dtype=np.float64).astype(np.float32)
    if prec == 1:
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ne,npb", [(36, 2), (7, 3), (24, 2), (32, 1), (48, 1)])
def test_rxchfne_one_pass(torch_dev, ne, npb):
    """neorx_rxchfne_omg2_dev: RXCHFne_FAE/FP/FBE_OMG2 from one pass over GM2
    stored in FP64, FP32 and BF16, against the oracle's three routines over
    the same rounded tensor.  The chunk widths differ: nebf 36 takes two
    chunks of 12 columns per lane (the second partial) and many row splits,
    nebf 7 one partial chunk of 8, nebf 24 one chunk of 9, nebf 32 one of 16
    with a single nuclear function, nebf 48 (rows of 2304 entries, the C3
    size) three chunks of 3 loads of four BF16 columns per lane.  Bars: 1e-11 for FP64 and FP32 storage
    (FP64 sums); for BF16, whose sums run in FP32, 1e-5 of the same
    contraction over absolute values, element by element."""
    torch = torch_dev
    rng = np.random.default_rng(ne)
    ng2 = ne ** 4 * npb ** 2
    G = rng.standard_normal(ng2) * 0.1
    DA, DB, DP = synth.density(ne, 20.0, 1), synth.density(ne, 2.0, 2), synth.density(npb, 1.0, 3)
    dG = _dev(torch, G)
    dA, dB, dP = (_dev(torch, x.ravel(order="F")) for x in (DA, DB, DP))
    F0 = [np.asfortranarray(rng.standard_normal((n, n))) for n in (ne, npb, ne)]
    for dtype in (0, 1, 2):
        print(f"case rxchfne one pass nebf={ne} npbf={npb} storage={dtype}", flush=True)
        if dtype == 0:
            store, Gr = dG, G
        elif dtype == 1:
            store = torch.empty(ng2, dtype=torch.float32, device="cuda")
            neorx.call_dev("neorx_convert_f32_dev", ng2, dG.data_ptr(), store.data_ptr(), 0)
            Gr = _round_storage(G, 1)
        else:
            store = torch.empty(ng2, dtype=torch.int16, device="cuda")
            neorx.call_dev("neorx_convert_bf16_dev", ng2, dG.data_ptr(), store.data_ptr(), 0)
            Gr = _round_storage(G, 2)
        dF = [_dev(torch, f.ravel(order="F")) for f in F0]
        dE = torch.zeros(1, dtype=torch.float64, device="cuda")
        neorx.call_dev("neorx_rxchfne_omg2_dev", ne, npb, dA.data_ptr(), dB.data_ptr(), dP.data_ptr(),
                       store.data_ptr(), dtype, *[f.data_ptr() for f in dF], dE.data_ptr(), 0)
        torch.cuda.synchronize()
        for k in range(3):
            n = npb if k == 1 else ne
            F = F0[k].copy(order="F")
            e = oracle.rxchfne_omg2(k, ne, npb, DA, DB, DP, Gr, F)
            got = dF[k].cpu().numpy().reshape(F.shape, order="F")
            if dtype < 2:
                check_close(got - F0[k], F - F0[k], what=f"RXCHF-ne one pass {k} storage {dtype}")
                assert dE.item() == pytest.approx(e, rel=1e-11), (dtype, k, dE.item(), e)
                continue
            # BF16: sums in FP32, whose error scales with the sum of the
            # terms' magnitudes (the same contraction over |D| and |G|), not
            # with the signed result where the terms cancel
            Fa = np.zeros((n, n), order="F")
            ea = oracle.rxchfne_omg2(k, ne, npb, np.abs(DA), np.abs(DB), np.abs(DP), np.abs(Gr), Fa)
            err = np.abs((got - F0[k]) - (F - F0[k]))
            assert np.all(err <= 1e-5 * Fa), (k, float((err / Fa).max()))
            assert abs(dE.item() - e) <= 1e-5 * ea, (k, dE.item(), e, ea)


def _run_list3(torch, S, cls, idx):
    n = len(idx)
    d_idx = torch.from_numpy(np.ascontiguousarray(idx)).cuda()
    out = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    S.prim_list3_dev(cls, n, d_idx.data_ptr(), *[o.data_ptr() for o in out],
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def test_rxchfmult(torch_dev):
    """RXCHFmult: class-5/6 tuples of a two-basis handle, RXCHFmult_GAM2,
    _GAM3_INT, _GAM4_INT and the FAE/FP/FBE_OMG{2,3,4} contractions through
    the C ABI against the oracle."""
    from _util import mult_tuples

    torch = torch_dev
    s = synth.mult(seed=51, e_shells=(0, 1, 2), be_shells=(0, 1, 2), p_shells=(0, 1, 2), be_centres=2)
    cs = synth.combined(s)
    S = neorx.System(s)
    for cls, n in ((5, 512), (6, 256)):
        print(f"case rxchfmult prim_list3 cls={cls} n={n}", flush=True)
        idx = mult_tuples(s, cls, n, seed=cls)
        got = _run_list3(torch, S, cls, idx)
        ref = oracle.prim_list3(cs, cls, idx)
        for k, what in enumerate(("ansE1", "ansS", "ansE2") if cls == 5 else ("ans",)):
            check_close(got[k], ref[k], what=f"cls{cls} {what}")
    S.close()
    t = synth.mult(seed=52, e_shells=(0, 1), be_shells=(0, 1), p_shells=(0, 0), n_e_centres=1, be_centres=1,
                   contract=(1, 2), ngtg=2)
    na, nb, npb = t["nebf"], t["nebfBE"], t["npbf"]
    print(f"case rxchfmult builders na={na} nb={nb} npbf={npb}", flush=True)
    T = neorx.System(t)
    got = T.RXCHFmult_GAM2()
    ref = oracle.rxchfmult_gam2(t)
    for g, r, what in zip(got, ref, ("INT_GAM2", "XCHF_GAM2", "XCHF_GAM2s")):
        check_close(g, r, what=what)
    g3 = T.RXCHFmult_GAM3_INT()
    r3 = oracle.rxchfmult_gam3_int(t)
    check_close(g3, r3, what="INT_GAM3")
    T.close()
    gee = oracle.gam_ee(t)
    g4 = neorx.RXCHFmult_GAM4_INT(na, nb, npb, ref[2], gee)
    r4 = oracle.rxchfmult_gam4_int(na, nb, npb, ref[2], gee)
    check_close(g4, r4, what="INT_GAM4")
    rng = np.random.default_rng(14)
    DA, DB, DP = synth.density(na, 2.0, 1), synth.density(nb, 1.0, 2), synth.density(npb, 1.0, 3)
    for order, G in ((2, ref[0]), (3, r3), (4, r4)):
        for which in (0, 1, 2):
            print(f"case rxchfmult omg{order} which={which}", flush=True)
            n = (na, npb, nb)[which]
            F0 = np.asfortranarray(rng.standard_normal((n, n)))
            A, B = F0.copy(order="F"), F0.copy(order="F")
            eg = neorx.RXCHFmult_OMG(which, order, na, nb, npb, DA, DB, DP, G, A)
            eo = oracle.rxchfmult_omg(order, which, na, nb, npb, DA, DB, DP, G, B)
            check_close(A - F0, B - F0, what=f"RXCHFmult omg{order} {which}")
            assert eg == pytest.approx(eo, rel=1e-11, abs=1e-14)
            # fixed-order sums: a second call agrees bit for bit
            A2 = F0.copy(order="F")
            assert neorx.RXCHFmult_OMG(which, order, na, nb, npb, DA, DB, DP, G, A2) == eg
            assert np.array_equal(A2, A)


def test_rxchf(torch_dev):
    """RXCHF with one special electron (source/rxchf/): class-9 tuples
    (RXCHF_xcalc_GAM2_MD), RXCHF_GAM2_IC1 and RXCHF_FAE/FP/FBE_OMG2 through
    the C ABI against the oracle (itself checked against the reference's
    routines in tests/test_oracle.py)."""
    torch = torch_dev
    s = synth.small(seed=150, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    S = neorx.System(s)
    idx = random_tuples(s, 2, 1024, seed=19)
    print("case rxchf cls9 n=1024", flush=True)
    got = _run_list3(torch, S, 9, idx)
    ref = oracle.prim_list3(s, 9, idx)
    for k, what in enumerate(("ansE1", "ansS", "ansE2")):
        check_close(got[k], ref[k], what=f"cls9 {what}")
    S.close()
    t = synth.small(seed=151, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    ne, npb = t["nebf"], t["npbf"]
    print(f"case rxchf builder nebf={ne} npbf={npb}", flush=True)
    T = neorx.System(t)
    g = T.RXCHF_GAM2_IC1()
    T.close()
    r = oracle.rxchf_gam2_ic1(t)
    for a, b, what in zip(g, r, ("GM2_1ICR", "GM2_2ICR", "GM2sICR")):
        check_close(a, b, what=f"RXCHF_GAM2_IC1 {what}")
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    rng = np.random.default_rng(23)
    for w in range(3):
        print(f"case rxchf omg2 which={w}", flush=True)
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        S0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B, C2, D = F0.copy(order="F"), F0.copy(order="F"), S0.copy(order="F"), S0.copy(order="F")
        eg = neorx.RXCHF_OMG2(w, ne, npb, DA, DB, DP, *g, A, C2)
        eo = oracle.rxchf_omg2(w, ne, npb, DA, DB, DP, *r, B, D)
        check_close(A - F0, B - F0, what=f"RXCHF OMG2 {w} F")
        check_close(C2 - S0, D - S0, what=f"RXCHF OMG2 {w} S")
        assert eg[0] == pytest.approx(eo[0], rel=1e-11) and eg[1] == pytest.approx(eo[1], rel=1e-11), (eg, eo)
    neorx.release_resident()


def test_rxchf3(torch_dev):
    """RXCHF e-e-e-p side (source/rxchf/): class-10 tuples
    (RXCHF_xcalc_GAM3_MD), RXCHF_GAM3_IC1, RXCHF_FAE/FP/FBE_OMG3 and
    RXCHF_FP/FBE_OMG1 through the C ABI against the oracle (itself checked
    against the reference's routines in tests/test_oracle.py); the
    contractions over random tensors, so that every index permutation of
    their formulas reads distinct values."""
    torch = torch_dev
    s = synth.small(seed=152, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    S = neorx.System(s)
    idx = random_tuples(s, 3, 512, seed=21)
    print("case rxchf3 cls10 n=512", flush=True)
    got = _run_list3(torch, S, 10, idx)
    ref = oracle.prim_list3(s, 10, idx)
    check_close(got[0], ref[0], what="cls10 ans1")
    check_close(got[1], ref[1], what="cls10 ans2")
    S.close()
    t = synth.small(seed=153, n_e_centres=3, e_shells=(0,), contract=(1, 2), p_shells=(0, 1), ngtg=2)
    print(f"case rxchf3 builder nebf={t['nebf']} npbf={t['npbf']}", flush=True)
    T = neorx.System(t)
    g = T.RXCHF_GAM3_IC1()
    T.close()
    r = oracle.rxchf_gam3_ic1(t)
    check_close(g[0], r[0], what="RXCHF_GAM3_IC1 GM3_1ICR")
    check_close(g[1], r[1], what="RXCHF_GAM3_IC1 GM3_2ICR")
    ne, npb = 4, 2
    rng = np.random.default_rng(29)
    G1, G2 = rng.standard_normal(ne ** 6 * npb * npb), rng.standard_normal(ne ** 6 * npb * npb)
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    for w in range(3):
        print(f"case rxchf3 omg3 which={w}", flush=True)
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        eg = neorx.RXCHF_OMG3(w, ne, npb, DA, DB, DP, G1, G2, A)
        eo = oracle.rxchf_omg3(w, ne, npb, DA, DB, DP, G1, G2, B)
        check_close(A - F0, B - F0, what=f"RXCHF OMG3 {w} F")
        assert eg == pytest.approx(eo, rel=1e-11), (w, eg, eo)
    H1, H1s = rng.standard_normal(ne * ne * npb * npb), rng.standard_normal(ne * ne * npb * npb)
    for w in (1, 2):
        print(f"case rxchf omg1 which={w}", flush=True)
        n = npb if w == 1 else ne
        F0 = np.asfortranarray(rng.standard_normal((n, n)))
        A, B = F0.copy(order="F"), F0.copy(order="F")
        SA, SB = np.asfortranarray(rng.standard_normal((n, n))), np.zeros((n, n), order="F")
        eg = neorx.RXCHF_OMG1(w, ne, npb, DB, DP, H1, H1s, A, SA)
        eo = oracle.rxchf_omg1(w, ne, npb, DB, DP, H1, H1s, B, SB)
        check_close(A - F0, B - F0, what=f"RXCHF OMG1 {w} F")
        check_close(SA, SB, what=f"RXCHF OMG1 {w} S")
        assert eg[0] == pytest.approx(eo[0], rel=1e-11) and eg[1] == pytest.approx(eo[1], rel=1e-11)
    neorx.release_resident()


def test_rxcuhf(torch_dev):
    """RXCUHF (open-shell regular electrons, source/rxchf/rxcuhf_omg{2,3}.f):
    RXCUHF_FE/FP/FBE_OMG2 and _OMG3 through the C ABI against the oracle
    (itself checked against the reference's routines in tests/test_oracle.py),
    over random tensors with different alpha and beta densities."""
    ne, npb = 4, 2
    rng = np.random.default_rng(31)
    Dal, Dbt, DB, DP = (synth.density(n, t, sd) for n, t, sd in ((ne, 2.0, 11), (ne, 1.0, 12), (ne, 1.0, 13),
                                                                  (npb, 1.0, 14)))
    Dto = Dal + Dbt
    G2a = [rng.standard_normal(ne ** 4 * npb * npb) for _ in range(3)]
    G3a = [rng.standard_normal(ne ** 6 * npb * npb) for _ in range(2)]
    for w in range(3):
        print(f"case rxcuhf omg2/omg3 which={w}", flush=True)
        n = npb if w == 1 else ne
        F0, S0, B0 = (np.asfortranarray(rng.standard_normal((n, n))) for _ in range(3))
        A, SA, BA = F0.copy(order="F"), S0.copy(order="F"), B0.copy(order="F")
        R, SR, BR = F0.copy(order="F"), S0.copy(order="F"), B0.copy(order="F")
        eg = neorx.RXCUHF_OMG2(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G2a, A, SA, BA)
        eo = oracle.rxcuhf_omg2(w, ne, npb, Dal, Dto, DB, DP, *G2a, R, SR, BR)
        check_close(A - F0, R - F0, what=f"RXCUHF OMG2 {w} F")
        check_close(SA - S0, SR - S0, what=f"RXCUHF OMG2 {w} S")
        if w == 0:
            check_close(BA - B0, BR - B0, what="RXCUHF OMG2 FAbetE")
        assert eg[0] == pytest.approx(eo[0], rel=1e-11) and eg[1] == pytest.approx(eo[1], rel=1e-11)
        A, BA, R, BR = F0.copy(order="F"), B0.copy(order="F"), F0.copy(order="F"), B0.copy(order="F")
        eg = neorx.RXCUHF_OMG3(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G3a, A, BA)
        eo = oracle.rxcuhf_omg3(w, ne, npb, Dal, Dbt, Dto, DB, DP, *G3a, R, BR)
        check_close(A - F0, R - F0, what=f"RXCUHF OMG3 {w} F")
        if w == 0:
            check_close(BA - B0, BR - B0, what="RXCUHF OMG3 FAbetE")
        assert eg == pytest.approx(eo, rel=1e-11)
    neorx.release_resident()


def test_rxchfmult_xchf_side(torch_dev):
    """RXCHFmult's XCHF-side builders over one basis: class-8 and class-7
    tuples, RXCHFmult_GAM1_MD, RXCHFmult_GAM3_XCHF and RXCHFmult_GAM4_XCHF."""
    from itertools import permutations

    torch = torch_dev
    s = synth.small(seed=61, e_shells=(0, 1, 2), p_shells=(0, 1, 2))
    S = neorx.System(s)
    idx = random_tuples(s, 1, 1024, seed=31)
    print("case rxchfmult cls8 n=1024", flush=True)
    got = _run_list3(torch, S, 8, idx)
    ref = oracle.prim_list3(s, 8, idx)
    check_close(got[0], ref[0], what="cls8 E")
    check_close(got[1], ref[1], what="cls8 S")
    base = random_tuples(s, 3, 40, seed=32)
    rows = [[x[p[0]], x[p[1]], x[p[2]], x[3], x[4 + p[0]], x[4 + p[1]], x[4 + p[2]], x[7]]
            for x in base for p in permutations(range(3))]
    idx = np.array(rows, dtype=np.int64)
    print("case rxchfmult cls7 n=240", flush=True)
    got = _run_list3(torch, S, 7, idx)[0].reshape(-1, 6).mean(axis=1)
    ref = oracle.prim_list3(s, 7, idx)[0].reshape(-1, 6).mean(axis=1)
    check_close(got, ref, what="cls7 six-permutation average")
    S.close()
    t = synth.small(seed=62, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 0), ngtg=2)
    print("case rxchfmult GAM1_MD / GAM3_XCHF", flush=True)
    T = neorx.System(t)
    g, gs = T.RXCHFmult_GAM1_MD()
    og, ogs = oracle.rxchfmult_gam1(t)
    check_close(g, og, what="RXCHFmult_GAM1_MD GAM_1")
    check_close(gs, ogs, what="RXCHFmult_GAM1_MD GAM_1s")
    check_close(T.RXCHFmult_GAM3_XCHF(), oracle.rxchfmult_gam3_xchf(t), what="RXCHFmult_GAM3_XCHF")
    T.close()
    reg = synth.small(seed=63, n_e_centres=1, e_shells=(0, 1, 0), contract=(1,), p_shells=(0,), ngtg=2)
    na, nb, npb = reg["nebf"], 4, reg["npbf"]
    gee = oracle.gam_ee(reg)
    g2s = np.random.default_rng(3).standard_normal(npb * npb * nb ** 4)
    eeb = np.ascontiguousarray(gee.reshape((na,) * 4)[:nb, :nb, :nb, :nb]).ravel()
    check_close(neorx.RXCHFmult_GAM4_XCHF(na, nb, npb, g2s, gee), oracle.gam4_icr(nb, npb, g2s, eeb),
                what="RXCHFmult_GAM4_XCHF")


def test_rxchfmult_exchange(torch_dev):
    """RXCHFmult with exchange (one basis): RXCHFmult_GAM2_IC1ex,
    RXCHFmult_GAM3_IC1ex and the FAE/FP/FBE_OMG2ex/_OMG3ex contractions."""
    s = synth.small(seed=66, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 0), ngtg=2)
    m = dict(s)
    m["nebfBE"], m["npebfBE"] = s["nebf"], s["npebf"]
    for k in synth.BE_KEYS:
        m[k + "_be"] = s[k]
    print("case rxchfmult GAM2_IC1ex / GAM3_IC1ex", flush=True)
    M = neorx.System(m)
    got2, ref2 = M.RXCHFmult_GAM2_IC1ex(), oracle.rxchfmult_gam2_ic1ex(s)
    for g, r, what in zip(got2, ref2, ("GM2_1ICR", "GM2_2ICR", "GM2_3ICR", "GM2sICR")):
        check_close(g, r, what=f"RXCHFmult_GAM2_IC1ex {what}")
    got3, ref3 = M.RXCHFmult_GAM3_IC1ex(), oracle.rxchfmult_gam3_ic1ex(s)
    for g, r, what in zip(got3, ref3, ("INT_GAM3", "INT_GAM3ex1", "INT_GAM3ex2")):
        check_close(g, r, what=f"RXCHFmult_GAM3_IC1ex {what}")
    M.close()
    ne, npb = s["nebf"], s["npbf"]
    rng = np.random.default_rng(15)
    DA, DB, DP = synth.density(ne, 2.0, 1), synth.density(ne, 1.0, 2), synth.density(npb, 1.0, 3)
    DA[0, 1] += 0.05
    for order, (G, X1, X2) in ((2, (ref2[1], ref2[2], None)), (3, ref3)):
        for which in (0, 1, 2):
            print(f"case rxchfmult omg{order}ex which={which}", flush=True)
            n = npb if which == 1 else ne
            F0 = np.asfortranarray(rng.standard_normal((n, n)))
            A, B = F0.copy(order="F"), F0.copy(order="F")
            eg = neorx.RXCHFmult_OMGex(which, order, ne, npb, DA, DB, DP, G, X1, X2, A)
            eo = oracle.rxchfmult_omgex(order, which, ne, npb, DA, DB, DP, G, X1, X2, B)
            check_close(A - F0, B - F0, what=f"RXCHFmult omg{order}ex {which}")
            assert eg == pytest.approx(eo, rel=1e-11, abs=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("cls", [1, 2, 3, 4])
def test_prim_tiers_lone_angular_momentum(torch_dev, cls):
    """The device's tier selection on tuples with all angular momentum on one
    particle (the capacity boundary of the one-particle kinetic terms) and on
    a larger random sample, against the oracle."""
    from test_engine_host import _lone_am_tuples

    print(f"case lone-am tiers cls={cls}", flush=True)
    s = synth.config2()
    S = neorx.System(s)
    ne = {1: 1, 2: 2, 3: 3, 4: 2}[cls]
    idx = np.concatenate([_lone_am_tuples(s, cls), random_tuples(s, ne, 512 if cls != 3 else 384, seed=700 + cls)])
    ge, _ = _run_list(torch_dev, S, cls, idx)
    oe, _ = oracle.prim_list(s, cls, idx)
    check_close(ge, oe, what=f"lone-am cls{cls}")
    S.close()


def test_rxchfmult_tiers_lone_angular_momentum(torch_dev):
    """RXCHFmult's XCHF-side tuples (classes 8 and 7) through the device's
    tier selection, on tuples with all angular momentum on one particle:
    class 8 against the oracle, class 7 after the six-permutation average
    (which moves the heavy electron through every position, e3 included)."""
    from itertools import permutations

    from test_engine_host import _lone_am_tuples

    torch = torch_dev
    s = synth.config2()
    S = neorx.System(s)
    idx = _lone_am_tuples(s, 8)
    print(f"case rxchfmult lone-am cls8 n={len(idx)}", flush=True)
    got = _run_list3(torch, S, 8, idx)
    ref = oracle.prim_list3(s, 8, idx)
    check_close(got[0], ref[0], what="lone-am cls8 E")
    base = _lone_am_tuples(s, 7)
    rows = [[x[p[0]], x[p[1]], x[p[2]], x[3], x[4 + p[0]], x[4 + p[1]], x[4 + p[2]], x[7]]
            for x in base for p in permutations(range(3))]
    idx = np.array(rows, dtype=np.int64)
    print(f"case rxchfmult lone-am cls7 n={len(idx)}", flush=True)
    got = _run_list3(torch, S, 7, idx)[0].reshape(-1, 6).mean(axis=1)
    ref = oracle.prim_list3(s, 7, idx)[0].reshape(-1, 6).mean(axis=1)
    check_close(got, ref, what="lone-am cls7 six-permutation average")
    S.close()


def test_uhf_contractions_reproducible(torch_dev):
    """FAE_OMG2/3 and UFP_OMG2/3 at nebf 8 (slab rows longer than a
    workgroup, so the row sums cross waves): equal to the oracle, and two
    calls on the same inputs agree bit for bit (fixed-order sums)."""
    print("case uhf reproducible nebf=8 npbf=2", flush=True)
    rng = np.random.default_rng(88)
    ne, npb, nae, nbe = 8, 2, 3, 2
    DAE, DBE = synth.density(ne, float(nae), 15), synth.density(ne, float(nbe), 16)
    DP = synth.density(npb, 1.0, 17)
    g2, g2s = rng.standard_normal(ne ** 4 * npb * npb), rng.standard_normal(ne ** 4 * npb * npb)
    g3 = rng.standard_normal(ne ** 6 * npb * npb)
    a = (nae, nbe, ne, npb, DAE, DBE, DP)
    for gfn, ofn, n in ((neorx.FAE_OMG2, oracle.fae_omg2, ne), (neorx.UFP_OMG2, oracle.ufp_omg2, npb)):
        runs = []
        for _ in range(2):
            Fg = np.zeros((n, n), order="F")
            Sg, eg, sg = gfn(*a, g2, g2s, Fg)
            runs.append((Fg, Sg, eg, sg))
        Fo, So = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
        eo, so = ofn(*a, g2, g2s, Fo, So)
        check_close(runs[0][0], Fo, what=gfn.__name__)
        check_close(runs[0][1], So, what=gfn.__name__ + " S")
        assert runs[0][2] == pytest.approx(eo, rel=1e-11, abs=1e-14)
        assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
        assert runs[0][2] == runs[1][2] and runs[0][3] == runs[1][3]
    for gfn, ofn, n in ((neorx.FAE_OMG3, oracle.fae_omg3, ne), (neorx.UFP_OMG3, oracle.ufp_omg3, npb)):
        runs = []
        for _ in range(2):
            Fg = np.zeros((n, n), order="F")
            runs.append((Fg, gfn(*a, g3, Fg)))
        Fo = np.zeros((n, n), order="F")
        eo = ofn(*a, g3, Fo)
        check_close(runs[0][0], Fo, what=gfn.__name__)
        assert runs[0][1] == pytest.approx(eo, rel=1e-11, abs=1e-14)
        assert np.array_equal(runs[0][0], runs[1][0]) and runs[0][1] == runs[1][1]
