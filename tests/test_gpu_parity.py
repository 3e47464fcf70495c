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


def test_gam3_builder(torch_dev):
    s = synth.small(seed=72, n_e_centres=3, e_shells=(0,), contract=(1,), p_shells=(0, 1))
    S = neorx.System(s)
    g3 = S.GAM3_IC1()
    o3 = oracle.gam3_ic1(s)
    check_close(g3, o3, what="GM3ICR")
    S.close()


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
