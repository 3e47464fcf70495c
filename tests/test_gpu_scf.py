"""GPU parity of the device-resident NEO-XCHF RHF path (nrx_scf.hip) through
the C ABI: the SCF's integrals, the Fock build of xchf1_fock with the
factorized 5-particle term, ROOTHAN, and whole SCF runs against the energies
of the reference's own xcscf (tests/golden/xchf_scf.json, made by
tests/golden/make_xchf_scf.py from the reference's sources)."""
import numpy as np
import pytest

import neorx
import oracle
from neorx import synth, xchf
from _util import check_close, golden_scf, golden_system

pytestmark = pytest.mark.gpu

# the golden values are printed with 10 decimals
PRINT_TOL = 1.5e-10


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    assert torch.cuda.is_available()
    torch.cuda.init()
    arch = neorx.init(0)
    assert arch.startswith("gfx950"), arch
    return torch


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _round_storage(g, prec):
    """GAM_3 as the device stores it in FP32 or BF16, widened back to FP64."""
    f = np.asarray(g, dtype=np.float64).astype(np.float32)
    if prec == "f32":
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _small(nelec, **kw):
    s = synth.small(**kw)
    s["nelec"] = nelec
    return s


def test_xchf_integrals(torch_dev):
    """class_nuc_rep, elec/nuc_ovlap, calc_GAM_epcore, calc_GAM_ep, calc_GAM_ee
    on the device against the oracle (p and d functions on both particles)."""
    s = _small(4, seed=81, n_e_centres=2, e_shells=(0, 1, 2), contract=(1, 2), p_shells=(0, 1, 2))
    print(f"case xchf integrals nebf={s['nebf']} npbf={s['npbf']}", flush=True)
    S = neorx.System(s)
    X = xchf.XCHF(S, 2)
    enuc, Se, Sp, He, Hp, Gep, Gee = X.integrals()
    oHe, oSe, oHp, oSp = oracle.one_body(s)
    assert enuc == pytest.approx(oracle.enuc(s), rel=1e-14)
    for g, o, nm in ((Se, oSe, "S_e"), (Sp, oSp, "S_p"), (He, oHe, "GAM_ecore"), (Hp, oHp, "GAM_pcore"),
                     (Gep, oracle.gam_ep(s), "GAM_ep"), (Gee, oracle.gam_ee(s), "GAM_ee")):
        check_close(g, o, what=nm)
    X.close()
    S.close()


@pytest.mark.parametrize("kw,sym", [
    (dict(seed=82, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0,)), True),
    (dict(seed=83, n_e_centres=3, e_shells=(0,), contract=(1, 2), p_shells=(0, 1)), False),
    (dict(seed=84, n_e_centres=1, e_shells=(0, 2), contract=(1, 1), p_shells=(0,), ngtg=2), True),
], ids=["sp-symmetric-D", "s-nonsymmetric-D", "sd-electrons"])
def test_xchf_fock(torch_dev, kw, sym):
    """One Fock build (xchf1_fock + XCfock_correction) against the oracle's
    restatement, which contracts the in-core GAM_4 of GAM4_ICR as EG4ICR does:
    checks the factorized 5-particle term, the shared-DE phase split and the
    assembly.  Also with a non-symmetric DE (the factorization does not
    assume symmetry)."""
    torch = torch_dev
    s = _small(4, **kw)
    ne, npb = s["nebf"], s["npbf"]
    print(f"case xchf fock nebf={ne} npbf={npb} sym={sym}", flush=True)
    rng = np.random.default_rng(ne)
    DE = synth.density(ne, 4.0, 3)
    if not sym:
        DE = np.asfortranarray(DE + 0.05 * rng.standard_normal((ne, ne)))
    DP = synth.density(npb, 1.0, 4)
    ints = oracle.xchf_ints(s)
    oFE, oFP, oc = oracle.xchf_fock(ne, npb, 4, ints, DE, DP)
    S = neorx.System(s)
    X = xchf.XCHF(S, 4, store=("f64", "f32", "bf16"))
    d_de, d_dp = _dev(torch, DE.ravel(order="F")), _dev(torch, DP.ravel(order="F"))
    fe = torch.zeros(ne * ne, dtype=torch.float64, device="cuda")
    fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
    c = torch.zeros(14, dtype=torch.float64, device="cuda")
    X.fock_dev(d_de.data_ptr(), d_dp.data_ptr(), "f64", fe.data_ptr(), fp.data_ptr(), c.data_ptr())
    torch.cuda.synchronize()
    gc = c.cpu().numpy()
    for k, nm in enumerate(xchf.COMPS):
        print(f"  {nm:9s} gpu {gc[k]: .15e} oracle {oc[k]: .15e} diff {gc[k] - oc[k]: .3e}")
    for k, nm in enumerate(xchf.COMPS):
        assert gc[k] == pytest.approx(oc[k], rel=1e-11, abs=1e-12 * abs(oc[0])), nm
    check_close(fe.cpu().numpy(), oFE.ravel(order="F"), what="FE")
    check_close(fp.cpu().numpy(), oFP.ravel(order="F"), what="FP")
    # reduced-precision storage of GAM_3 moves only E_GAM3 and the matrices,
    # by the storage rounding; the build equals the oracle's FP64 build over
    # the same rounded GAM_3 (FP32: round to nearest; BF16: the FP32 value
    # rounded to nearest even in its upper 16 bits) -- FP32 storage to the
    # FP64 bar (its sums run in FP64), BF16 storage to 1e-6 (its sums over 24
    # to 73 products run in FP32, 2^-24 per operation, far below its storage
    # rounding of 2^-9)
    for prec, tol in (("f32", 1e-6), ("bf16", 1e-2)):
        bar = dict(rtol=1e-11, atol_rel=1e-12) if prec == "f32" else dict(rtol=1e-6, atol_rel=1e-6)
        X.fock_dev(d_de.data_ptr(), d_dp.data_ptr(), prec, fe.data_ptr(), fp.data_ptr(), c.data_ptr())
        torch.cuda.synchronize()
        gl = c.cpu().numpy()
        assert abs(gl[8] - oc[8]) <= tol * abs(oc[8]), (prec, gl[8], oc[8])
        assert gl[7] == pytest.approx(oc[7], rel=1e-11) and gl[9] == pytest.approx(oc[9], rel=1e-11)
        rints = dict(ints, G3=_round_storage(ints["G3"], prec))
        rFE, rFP, rc = oracle.xchf_fock(ne, npb, 4, rints, DE, DP)
        for k, nm in enumerate(xchf.COMPS):
            assert gl[k] == pytest.approx(rc[k], rel=bar["rtol"], abs=bar["atol_rel"] * abs(rc[0])), (prec, nm)
        check_close(fe.cpu().numpy(), rFE.ravel(order="F"), what=f"FE {prec}", **bar)
        check_close(fp.cpu().numpy(), rFP.ravel(order="F"), what=f"FP {prec}", **bar)
    X.close()
    S.close()


def test_xchf_fock_incremental(torch_dev):
    """The incremental build: rows of a reference density D0 (an FP64 build
    kept as the reference) plus T(D - D0, D + D0) equal the full build at D.
    Read in FP64 this is the exact bilinear identity (parity with the oracle
    at D to 1e-11); read in BF16 the increment's storage error scales with
    |D - D0|, far below that of a full BF16 build."""
    torch = torch_dev
    s = _small(4, seed=85, n_e_centres=1, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1))
    ne, npb = s["nebf"], s["npbf"]
    print(f"case xchf fock incremental nebf={ne} npbf={npb}", flush=True)
    rng = np.random.default_rng(7)
    D0 = synth.density(ne, 4.0, 3)
    E = rng.standard_normal((ne, ne))
    D = np.asfortranarray(D0 + 1e-3 * (E + E.T))
    DP = synth.density(npb, 1.0, 4)
    oFE, oFP, oc = oracle.xchf_fock(ne, npb, 4, oracle.xchf_ints(s), D, DP)
    S = neorx.System(s)
    X = xchf.XCHF(S, 4, store=("f64", "bf16"))
    d0, d1, dp = (_dev(torch, a.ravel(order="F")) for a in (D0, D, DP))
    fe = torch.zeros(ne * ne, dtype=torch.float64, device="cuda")
    fp = torch.zeros(npb * npb, dtype=torch.float64, device="cuda")
    c = torch.zeros(14, dtype=torch.float64, device="cuda")
    X.fock_dev(d0.data_ptr(), dp.data_ptr(), "f64", fe.data_ptr(), fp.data_ptr(), c.data_ptr(), g3pass=1)
    X.fock_dev(d1.data_ptr(), dp.data_ptr(), "f64", fe.data_ptr(), fp.data_ptr(), c.data_ptr(), g3pass=2)
    torch.cuda.synchronize()
    gc = c.cpu().numpy()
    for k, nm in enumerate(xchf.COMPS):
        assert gc[k] == pytest.approx(oc[k], rel=1e-11, abs=1e-12 * abs(oc[0])), nm
    check_close(fe.cpu().numpy(), oFE.ravel(order="F"), what="FE")
    check_close(fp.cpu().numpy(), oFP.ravel(order="F"), what="FP")
    # BF16 increment vs a full BF16 build at D
    X.fock_dev(d1.data_ptr(), dp.data_ptr(), "bf16", fe.data_ptr(), fp.data_ptr(), c.data_ptr(), g3pass=2)
    torch.cuda.synchronize()
    e_inc = c.cpu().numpy()[8]
    X.fock_dev(d1.data_ptr(), dp.data_ptr(), "bf16", fe.data_ptr(), fp.data_ptr(), c.data_ptr())
    torch.cuda.synchronize()
    e_full = c.cpu().numpy()[8]
    err_inc, err_full = abs(e_inc - oc[8]), abs(e_full - oc[8])
    print(f"  E_GAM3 {oc[8]:.12e}: BF16 increment error {err_inc:.2e}, full BF16 build error {err_full:.2e}")
    assert err_inc <= 1e-2 * err_full + 1e-14 * abs(oc[8]), (err_inc, err_full)
    X.close()
    S.close()


@pytest.mark.parametrize("M,N,K,ta,tb", [(16, 16, 4, False, False), (24, 13824, 24, False, False),
                                         (24, 24, 13824, False, True), (5, 37, 9, True, False),
                                         (33, 7, 70, True, True)])
def test_gemm_f64(torch_dev, M, N, K, ta, tb):
    """The FP64 MFMA GEMM (v_mfma_f64_16x16x4_f64 lane layout, edge tiles,
    strided operands, split-K) against numpy."""
    torch = torch_dev
    rng = np.random.default_rng(M * 1000 + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    As = np.ascontiguousarray(A.T) if ta else A
    Bs = np.ascontiguousarray(B.T) if tb else B
    sam, sak = (1, M) if ta else (K, 1)
    sbk, sbn = (1, K) if tb else (N, 1)
    dA, dB = _dev(torch, As.ravel()), _dev(torch, Bs.ravel())
    dC = torch.zeros(M * N, dtype=torch.float64, device="cuda")
    neorx.call_dev("neorx_gemm_f64_dev", M, N, K, dA.data_ptr(), sam, sak, dB.data_ptr(), sbk, sbn, dC.data_ptr(), 0)
    torch.cuda.synchronize()
    ref = A @ B
    got = dC.cpu().numpy().reshape(M, N)
    assert np.max(np.abs(got - ref)) <= 1e-13 * K * np.max(np.abs(ref)), np.max(np.abs(got - ref))


@pytest.mark.parametrize("n", [1, 7, 24, 50, 70])
def test_roothan(torch_dev, n):
    """ROOTHAN on the device (Jacobi; S^-1/2 in LDS for n <= 62, the
    diagonalisation in LDS for n <= 43, global memory above)
    against the oracle's restatement: eigenvalues, the density, and the
    orbitals up to their sign; F C = S C e and C^T S C = 1."""
    torch = torch_dev
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    Sm = np.asfortranarray(A @ A.T / n + np.eye(n))
    Fm = rng.standard_normal((n, n))
    Fm = np.asfortranarray(Fm + Fm.T)
    k1, k2 = (1, max(1, n // 3)) if n > 1 else (1, 1)
    oC, oev, oD = oracle.roothan(Sm, Fm, k1, k2, 2.0)
    d = [torch.zeros(m, dtype=torch.float64, device="cuda") for m in (n * n, n, n * n)]
    dS, dF = _dev(torch, Sm.ravel(order="F")), _dev(torch, Fm.ravel(order="F"))  # alive until synchronised
    xchf.roothan_dev(n, dS.data_ptr(), dF.data_ptr(), k1, k2, 2.0, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
    torch.cuda.synchronize()
    C, ev, D = (x.cpu().numpy() for x in d)
    C = C.reshape(n, n, order="F")
    check_close(ev, oev, rtol=1e-12, atol_rel=1e-13, what="eigenvalues")
    check_close(D, oD.ravel(order="F"), rtol=1e-10, atol_rel=1e-12, what="density")
    sgn = np.sign(np.sum(C * oC, axis=0))
    check_close(C * sgn, oC, rtol=1e-9, atol_rel=1e-11, what="orbitals")
    assert np.max(np.abs(C.T @ Sm @ C - np.eye(n))) < 1e-12
    assert np.max(np.abs(Fm @ C - Sm @ C * ev)) < 1e-11 * np.max(np.abs(Fm))


def _run(name, **kw):
    g = golden_scf()[name]
    s, guess = golden_system(g)
    S = neorx.System(s)
    X = xchf.XCHF(S, s["nelec"], store=kw.pop("store", ("f64",)))
    r = X.scf(ce_guess=guess, nucst=s.get("nucst", 1), **kw)
    X.close()
    S.close()
    return g, r


def _check_golden(g, r, skip=()):
    E = r["comps"]["E_TOTAL"]
    print(f"  E = {E:.12f} (reference {g['energy']:.10f}) in {r['niter']} iterations", flush=True)
    assert abs(E - g["energy"]) <= 1e-8, (E, g["energy"])
    for k, v in g["initial"].items():
        if k not in skip:
            assert abs(r["comps_init"][k] - v) <= PRINT_TOL + 1e-12 * abs(v), (k, r["comps_init"][k], v)


@pytest.mark.parametrize("graph,diis", [(True, False), (True, True), (False, True)])
def test_scf_example(torch_dev, graph, diis):
    """examples/xchf of the reference (He2 with a quantum proton, guess orbitals
    from guessCE.inp): the initial-guess energy components and the final
    energy -5.7800665094 of examples/xchf/answer/XCHF.log:335, FP64."""
    print(f"case scf example graph={graph} diis={diis}", flush=True)
    g, r = _run("example", graph=graph, diis=diis)
    _check_golden(g, r)


@pytest.mark.parametrize("name", ["c1mini_sp", "tiny_s", "tiny_pnuc"])
def test_scf_golden_minis(torch_dev, name):
    """Small systems in the C1 geometry (s/p electrons, s/p nuclear
    functions), core guess: the reference xcscf's initial components and
    final energy (tiny_s and tiny_pnuc from its ROOTHAN steps)."""
    print(f"case scf {name}", flush=True)
    g, r = _run(name, diis=True)
    _check_golden(g, r)


@pytest.mark.parametrize("mode", ["f32", "bf16", "inc", "inc-f64ref"])
def test_scf_mixed_precision(torch_dev, mode):
    """Mixed precision: GAM_3 read in FP32 (or BF16, then FP32) until the DIIS
    error falls below 1e-5, then FP64; or (inc) BF16 builds, then BF16-stored
    increments from reference builds: the first reference from the FP32 copy
    when the context holds one (inc), from FP64 when it does not
    (inc-f64ref), every later one from FP64.  Every mode ends on an FP64
    build, so the final energy is the FP64 one."""
    print(f"case scf mixed {mode}", flush=True)
    store = ("f64", "bf16") if mode == "inc-f64ref" else ("f64", "f32", "bf16")
    g, r = _run("c1mini_sp", mode=mode.split("-")[0], diis=True, store=store)
    prec, kind = r["trace"][:, 5], r["trace"][:, 7]
    assert prec[0] == {"f32": 1, "bf16": 2, "inc": 2, "inc-f64ref": 2}[mode] and prec[-1] == 0, prec
    if mode.startswith("inc"):
        # BF16 builds first, increments in between, an FP64 reference last
        assert kind[0] == 0 and kind[-1] == 1 and (kind == 2).any(), kind
        refs = prec[kind == 1]
        assert np.all(prec[kind == 2] == 2) and refs[-1] == 0, (prec, kind)
        assert refs[0] == (1 if mode == "inc" else 0) and np.all(refs[1:] == 0), (prec, kind)
    # the first build reads GAM_3 in the reduced storage: E_GAM3 (and the
    # total) of the initial guess move by the storage rounding, the rest not
    tol = {"f32": 1e-6, "bf16": 1e-2, "inc": 1e-2, "inc-f64ref": 1e-2}[mode]
    assert abs(r["comps_init"]["E_GAM3"] - g["initial"]["E_GAM3"]) <= tol * abs(g["initial"]["E_GAM3"]) + PRINT_TOL
    _check_golden(g, r, skip=("E_GAM3", "E_TOTAL"))


def test_scf_graph_matches_eager(torch_dev):
    """The hipGraph replay of an iteration and the eager launches give the
    same energies bit for bit (every reduction has a fixed order)."""
    print("case scf graph vs eager c1mini_sp", flush=True)
    _, a = _run("c1mini_sp", graph=True, diis=True)
    _, b = _run("c1mini_sp", graph=False, diis=True)
    assert a["niter"] == b["niter"]
    assert np.array_equal(a["trace"][:, :5], b["trace"][:, :5])


def test_scf_batch_concurrent(torch_dev):
    """Many instances at once (reduced config 5: the proton scanned along the
    bond, uneven basis sizes): xchf.scf_batch runs each instance's SCF from
    its own host thread and stream, concurrently, graph captures included;
    every energy equals the
    one of the same instance run alone, bit for bit, in FP64 and in the
    incremental mixed mode, and the mixed energies are within 1e-6 Eh of
    the FP64 ones.  One small instance is also checked against the oracle's
    xcscf (ROOTHAN steps, no DIIS) to 1e-8 Eh."""
    from neorx import xchf

    ks = list(range(0, 64, 8))
    systems = [synth.config5(k) for k in ks]
    print(f"case scf batch {len(ks)} instances nebf {[s['nebf'] for s in systems]}", flush=True)
    hs = [neorx.System(s) for s in systems]
    xs = [xchf.XCHF(h, s["nelec"], store=("f64", "bf16")) for h, s in zip(hs, systems)]
    kw = dict(diis=True, graph=True, maxit=80, sw_hi=1e-3, sw_lo=1e-6)
    e = {}
    for mode in ("f64", "inc"):
        # concurrent first: each context captures its iteration graphs while
        # the other threads launch, copy and capture on their own streams
        con = [r["comps"]["E_TOTAL"] for r in xchf.scf_batch(xs, len(xs), mode=mode, **kw)]
        seq = [x.scf(mode=mode, **kw)["comps"]["E_TOTAL"] for x in xs]
        assert seq == con, (mode, seq, con)
        e[mode] = np.array(con)
    assert np.max(np.abs(e["inc"] - e["f64"])) <= 1e-6, e
    # one instance against the oracle's SCF (no DIIS, as xcscf)
    k = min(range(len(systems)), key=lambda i: systems[i]["nebf"] ** 6 * systems[i]["npbf"] ** 2)
    s = systems[k]
    print(f"case scf batch oracle instance nebf={s['nebf']} npbf={s['npbf']}", flush=True)
    r = xs[k].scf(mode="f64", diis=False, graph=True, maxit=200)
    o = oracle.xcscf(s, oracle.xchf_ints(s), maxit=200)
    assert r["comps"]["E_TOTAL"] == pytest.approx(o["energy"], abs=1e-8), (r["comps"]["E_TOTAL"], o["energy"])
    for x in xs:
        x.close()
    for h in hs:
        h.close()
