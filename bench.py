"""Benchmark of the geminal-integral hot path on one MI355X.

Workload (config "integral kernels only"): 60 electronic Cartesian
primitives (s/p/d) on 4 centres, 16 nuclear primitives (l<=2), 4-term
geminal, a Z=1 classical nucleus per centre (neorx.synth.config2).

One step = one pass of the primitive-integral kernels over
  * every e-p tuple (xcalc_GAM1_MD values, 921,600 tuples),
  * a seeded uniform sample of NS e-e-p tuples (xcalc_GAM2_MD values),
  * a seeded uniform sample of NS e-e-e-p tuples (xcalc_GAM3_MD values),
all inputs resident in HBM before the timed region.  value = tuples of the
step / step time (integrals/s, one integral = one fully geminal-summed
primitive tuple).  Per-class rates are reported alongside.

The cpu_baseline leg times the reference's own routines (oracle/_ref, built
from the reference sources) on a bounded sample of the same tuples using the
host's cores, or the oracle when the reference checker is absent.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neo-rxchf_amd"))

METRIC = "gam2-4 integrals/s + SCF s/iter: mixed vs FP64 GPU at |dE|<=1e-6 Eh"
FP64_PEAK_TFLOPS = 78.6


def sample_tuples(np, sysd, cls, n, seed):
    rng = np.random.default_rng(seed)
    w = 2 * (cls + 1)
    idx = np.zeros((n, w), dtype=np.int64)
    for i in list(range(cls)) + list(range(cls + 1, 2 * cls + 1)):
        idx[:, i] = rng.integers(0, sysd["npebf"], n)
    idx[:, cls] = rng.integers(0, sysd["npbf"], n)
    idx[:, 2 * cls + 1] = rng.integers(0, sysd["npbf"], n)
    return idx


def cpu_baseline(np, sysd, budget_s, cores):
    """Reference (or oracle) primitive-tuple rates on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from concurrent.futures import ThreadPoolExecutor

    kind = "reference" if oracle.ref_available() else "port"
    rates = {}
    detail = {}
    for cls in (1, 2, 3):
        per = budget_s / 3.0
        done = 0
        t0 = time.perf_counter()
        seed = 1000 + cls
        with ThreadPoolExecutor(max_workers=cores) as ex:
            while time.perf_counter() - t0 < per:
                batch = sample_tuples(np, sysd, cls, cores, seed)
                seed += 1
                if kind == "reference":
                    futs = [ex.submit(oracle.ref_prim_list, sysd, cls, batch[i:i + 1]) for i in range(cores)]
                else:
                    futs = [ex.submit(oracle.prim_list, sysd, cls, batch[i:i + 1]) for i in range(cores)]
                for f in futs:
                    f.result()
                done += cores
        dt = time.perf_counter() - t0
        rates[cls] = done / dt
        detail[cls] = (done, dt)
    return kind, rates, detail


def fock_bench(np, torch, neorx, synth, ne, npb, reps):
    """One in-core Fock build's contractions (EG2IC1 + EG3IC1) at nebf=ne,
    npbf=npb on synthetic tensors of that size, FP64 vs FP32-stored tensors,
    plus the energy difference of the two paths on a real small system."""
    f64, f32 = torch.float64, torch.float32
    ng2, ng3 = ne ** 4 * npb * npb, ne ** 6 * npb * npb
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    g2 = torch.randn(ng2, dtype=f64, device="cuda", generator=gen)
    g2s = torch.randn(ng2, dtype=f64, device="cuda", generator=gen)
    g3 = torch.randn(ng3, dtype=f64, device="cuda", generator=gen)
    h = {k: torch.empty(v.numel(), dtype=f32, device="cuda") for k, v in (("g2", g2), ("g2s", g2s), ("g3", g3))}
    for k, v in (("g2", g2), ("g2s", g2s), ("g3", g3)):
        neorx.call_dev("neorx_convert_f32_dev", v.numel(), v.data_ptr(), h[k].data_ptr(), 0)
    DE = torch.from_numpy(np.ascontiguousarray(synth.density(ne, 14.0, 1).ravel(order="F"))).cuda()
    DP = torch.from_numpy(np.ascontiguousarray(synth.density(npb, 1.0, 2).ravel(order="F"))).cuda()
    out = [torch.zeros(n, dtype=f64, device="cuda") for n in (ne * ne, ne * ne, npb * npb, npb * npb, 2, 1)]
    ptr = [o.data_ptr() for o in out]

    def build(mixed):
        if mixed:
            neorx.call_dev("neorx_eg2ic1_f32_dev", ne, npb, DE.data_ptr(), DP.data_ptr(), h["g2"].data_ptr(),
                           h["g2s"].data_ptr(), ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[4] + 8, 0)
            ev[0].record()
            neorx.call_dev("neorx_eg3ic1_f32_dev", ne, npb, DE.data_ptr(), DP.data_ptr(), h["g3"].data_ptr(), ptr[0],
                           ptr[2], ptr[5], 0)
            ev[1].record()
        else:
            neorx.call_dev("neorx_eg2ic1_dev", ne, npb, DE.data_ptr(), DP.data_ptr(), g2.data_ptr(), g2s.data_ptr(),
                           ptr[0], ptr[1], ptr[2], ptr[3], ptr[4], ptr[4] + 8, 0)
            ev[0].record()
            neorx.call_dev("neorx_eg3ic1_dev", ne, npb, DE.data_ptr(), DP.data_ptr(), g3.data_ptr(), ptr[0], ptr[2],
                           ptr[5], 0)
            ev[1].record()

    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    res = {}
    for mixed in (False, True):
        build(mixed)
        torch.cuda.synchronize()
        t, t3 = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            build(mixed)
            torch.cuda.synchronize()
            t.append((time.perf_counter() - t0) * 1e3)
            t3.append(ev[0].elapsed_time(ev[1]))
        res[mixed] = (min(t), min(t3))
    del g2, g2s, g3, h
    torch.cuda.empty_cache()
    # energy difference of the two paths on real integrals (small system)
    sysd = synth.small(seed=3, n_e_centres=2, e_shells=(0, 1), contract=(1, 2), p_shells=(0, 1), ngtg=3)
    S = neorx.System(sysd)
    n2, n3 = sysd["nebf"] ** 4 * sysd["npbf"] ** 2, sysd["nebf"] ** 6 * sysd["npbf"] ** 2
    G2, G2s, G3 = (torch.empty(n, dtype=f64, device="cuda") for n in (n2, n2, n3))
    neorx.call_dev("neorx_gam2_ic1_dev", S.handle, G2.data_ptr(), G2s.data_ptr(), 0)
    neorx.call_dev("neorx_gam3_ic1_dev", S.handle, G3.data_ptr(), 0)
    F2, F2s, F3 = (torch.empty(n, dtype=f32, device="cuda") for n in (n2, n2, n3))
    for a, b in ((G2, F2), (G2s, F2s), (G3, F3)):
        neorx.call_dev("neorx_convert_f32_dev", a.numel(), a.data_ptr(), b.data_ptr(), 0)
    me, mp = sysd["nebf"], sysd["npbf"]
    de = torch.from_numpy(np.ascontiguousarray(synth.density(me, 4.0, 3).ravel(order="F"))).cuda()
    dp = torch.from_numpy(np.ascontiguousarray(synth.density(mp, 1.0, 4).ravel(order="F"))).cuda()
    energies = {}
    for mixed in (False, True):
        o = [torch.zeros(n, dtype=f64, device="cuda") for n in (me * me, me * me, mp * mp, mp * mp, 2, 1)]
        q = [x.data_ptr() for x in o]
        if mixed:
            neorx.call_dev("neorx_eg2ic1_f32_dev", me, mp, de.data_ptr(), dp.data_ptr(), F2.data_ptr(),
                           F2s.data_ptr(), q[0], q[1], q[2], q[3], q[4], q[4] + 8, 0)
            neorx.call_dev("neorx_eg3ic1_f32_dev", me, mp, de.data_ptr(), dp.data_ptr(), F3.data_ptr(), q[0], q[2],
                           q[5], 0)
        else:
            neorx.call_dev("neorx_eg2ic1_dev", me, mp, de.data_ptr(), dp.data_ptr(), G2.data_ptr(), G2s.data_ptr(),
                           q[0], q[1], q[2], q[3], q[4], q[4] + 8, 0)
            neorx.call_dev("neorx_eg3ic1_dev", me, mp, de.data_ptr(), dp.data_ptr(), G3.data_ptr(), q[0], q[2], q[5],
                           0)
        torch.cuda.synchronize()
        energies[mixed] = float(o[4].cpu()[0] + o[5].cpu()[0])
    S.close()
    bytes3 = {False: 8 * ng3, True: 4 * ng3}
    return {
        "tensors": f"nebf={ne} npbf={npb}: GM2 {ng2} x2, GM3 {ng3} entries (synthetic values)",
        "ms_fp64": res[False][0], "ms_mixed": res[True][0], "speedup": res[False][0] / res[True][0],
        "eg3_ms_fp64": res[False][1], "eg3_ms_mixed": res[True][1],
        "eg3_GBps_fp64": bytes3[False] / res[False][1] / 1e6, "eg3_GBps_mixed": bytes3[True] / res[True][1] / 1e6,
        "dE_mixed_vs_fp64_Eh": abs(energies[True] - energies[False]),
        "dE_system": f"nebf={me} npbf={mp} real GAM2/GAM3 integrals, E_gam2 + E_gam3 = {energies[False]:.12f} Eh",
    }


def c3_bench(np, torch, neorx, synth, reps=5):
    """Config 3: RXCHF-ne closed shell with 2 special electrons (48/12
    functions, 3 geminals).  RXCHFne_GAM2_IC1 is built on the device in FP64;
    the three closed-shell contractions (RXCHFne_FAE/FP/FBE_OMG2, one pass over
    GM2) then read it stored in FP64, FP32 and BF16.  Reports the pass time per
    storage type and the energy and Fock-matrix error of the reduced-precision
    passes against the FP64 one, with the same densities (random PSD: DAE of
    trace 20, DBE of trace 2, DP of trace 1)."""
    f64 = torch.float64
    sysd = synth.config3()
    ne, npb = sysd["nebf"], sysd["npbf"]
    ng2 = ne ** 4 * npb ** 2
    S = neorx.System(sysd)
    g2 = torch.empty(ng2, dtype=f64, device="cuda")
    gx = torch.empty(ng2, dtype=f64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    neorx.call_dev("neorx_rxchfne_gam2_ic1_dev", S.handle, g2.data_ptr(), gx.data_ptr(), 0)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    S.close()
    del gx
    torch.cuda.empty_cache()
    store = {0: g2, 1: torch.empty(ng2, dtype=torch.float32, device="cuda"),
             2: torch.empty(ng2, dtype=torch.int16, device="cuda")}
    neorx.call_dev("neorx_convert_f32_dev", ng2, g2.data_ptr(), store[1].data_ptr(), 0)
    neorx.call_dev("neorx_convert_bf16_dev", ng2, g2.data_ptr(), store[2].data_ptr(), 0)
    dens = [torch.from_numpy(np.ascontiguousarray(synth.density(n, tr, sd).ravel(order="F"))).cuda()
            for n, tr, sd in ((ne, 20.0, 31), (ne, 2.0, 32), (npb, 1.0, 33))]
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    res = {}
    for dt in (0, 1, 2):
        outs = None
        ts = []
        for r in range(reps + 1):
            outs = [torch.zeros(n * n, dtype=f64, device="cuda") for n in (ne, npb, ne)] + \
                   [torch.zeros(1, dtype=f64, device="cuda")]
            ev[0].record()
            neorx.call_dev("neorx_rxchfne_omg2_dev", ne, npb, *[d.data_ptr() for d in dens], store[dt].data_ptr(),
                           dt, *[o.data_ptr() for o in outs], 0)
            ev[1].record()
            torch.cuda.synchronize()
            if r:
                ts.append(ev[0].elapsed_time(ev[1]))
        res[dt] = (min(ts), [o.cpu().numpy() for o in outs])
    out = {"system": sysd["name"], "gm2_entries": ng2, "gam2_ic1_build_s": t_build}
    ref = res[0][1]
    for dt, nm in ((0, "fp64"), (1, "fp32"), (2, "bf16")):
        ms, o = res[dt]
        out[f"pass_ms_{nm}"] = ms
        out[f"pass_GBps_{nm}"] = ng2 * (8, 4, 2)[dt] / ms / 1e6
        out[f"energy_{nm}_Eh"] = float(o[3][0])
        if dt:
            out[f"speedup_{nm}_vs_fp64"] = res[0][0] / ms
            out[f"dE_{nm}_vs_fp64_Eh"] = abs(float(o[3][0]) - float(ref[3][0]))
            out[f"dF_{nm}_vs_fp64_rel"] = max(float(np.abs(o[k] - ref[k]).max() / np.abs(ref[k]).max())
                                             for k in range(3))
    return out


def scf_bench(np, torch, neorx, synth, maxit=80):
    """NEO-XCHF RHF SCF of the C1 system (neorx.synth.config1: C/N classical,
    quantum proton, 24 electronic functions from 36 s/p primitives, 8 nuclear
    s/p/d primitives, 3-term geminal, 14 electrons) on the device: every
    integral built on the GPU (GM3 = 1.22e10 entries, kept in FP64, FP32 and
    BF16: 171 GB), then the SCF with DIIS, each iteration one replay of a
    captured hipGraph:
      * FP64: GAM_3 read in FP64 every iteration;
      * mixed (incremental): GAM_3 read in BF16 (sums over its stored
        entries in FP32) until the DIIS error < 1e-3, then a build from the
        FP32 copy kept as the reference and BF16-stored increments
        T(DE - DEref, DE + DEref) from it, an FP64 reference once the DIIS
        error reaches 1e-6, and a last FP64 build that confirms convergence.
    Each mode runs once untimed (graph capture) and once timed; SCF s/iter =
    the timed run's wall time (host clock, first graph launch to the last
    read-back) / its iterations.  The EG3 slab depth is autotuned per storage
    type on the built tensors."""
    from neorx import xchf

    sysd = synth.config1()
    S = neorx.System(sysd)
    t0 = time.perf_counter()
    X = xchf.XCHF(S, sysd["nelec"], store=("f64", "f32", "bf16"), autotune=True)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    ne, npe, npb = sysd["nebf"], sysd["npebf"], sysd["npbf"]
    prims = npe ** 6 * npb ** 2 + npe ** 4 * npb ** 2 + npe ** 2 * npb ** 2
    print(f"# C1 integrals built in {t_build:.1f} s", file=sys.stderr, flush=True)
    runs = {}
    names = {0: "f64", 1: "f32", 2: "bf16"}
    for mode in ("f64", "inc"):
        X.scf(mode=mode, diis=True, graph=True, maxit=maxit, sw_hi=1e-3, sw_lo=1e-6)  # captures the graphs
        t0 = time.perf_counter()
        r = X.scf(mode=mode, diis=True, graph=True, maxit=maxit, sw_hi=1e-3, sw_lo=1e-6)
        wall = time.perf_counter() - t0
        tr = r["trace"]
        kinds = {}
        for k in range(len(tr)):
            nm = names[int(tr[k, 5])] + ("-increment" if tr[k, 7] == 2 else "-reference" if tr[k, 7] == 1 else "")
            kinds.setdefault(nm, []).append(float(tr[k, 6]))
        per = {k: {"iterations": len(v), "ms_per_iter": float(np.median(v))} for k, v in kinds.items()}
        runs[mode] = {"energy": r["comps"]["E_TOTAL"], "iterations": r["niter"], "wall_s": wall,
                      "s_per_iter": wall / r["niter"], "by_build": per}
        print(f"# SCF {mode}: E={r['comps']['E_TOTAL']:.10f} it={r['niter']} {wall / r['niter'] * 1e3:.2f} ms/iter "
              + " ".join(f"{k}:{v['iterations']}x{v['ms_per_iter']:.2f}ms" for k, v in per.items()),
              file=sys.stderr, flush=True)
    X.close()
    S.close()
    f64, mix = runs["f64"], runs["inc"]
    return {
        "system": "C1 NEO-XCHF RHF: nebf 24 (36 s/p prims), npbf 8 (s/p/d), 3 geminals, 14 electrons",
        "integral_build_s": t_build, "primitive_tuples_built": prims,
        "scf_s_per_iter_fp64": f64["s_per_iter"], "scf_s_per_iter_mixed": mix["s_per_iter"],
        "speedup_s_per_iter_mixed_vs_fp64": f64["s_per_iter"] / mix["s_per_iter"],
        "speedup_bf16_increment_iteration_vs_fp64": (f64["by_build"]["f64"]["ms_per_iter"] /
                                                     mix["by_build"]["bf16-increment"]["ms_per_iter"]
                                                     if "bf16-increment" in mix["by_build"] else None),
        "scf_wall_s_fp64": f64["wall_s"], "scf_wall_s_mixed": mix["wall_s"],
        "speedup_scf_wall_mixed_vs_fp64": f64["wall_s"] / mix["wall_s"],
        "energy_fp64_Eh": f64["energy"], "dE_mixed_vs_fp64_Eh": abs(mix["energy"] - f64["energy"]),
        "runs": runs,
    }


def n5_bench(np, torch, neorx, synth, n=64, workers=16, maxit=80):
    """Config 5 at reduced basis sizes: n independent NEO-XCHF RHF instances
    (neorx.synth.config5: proton scanned over n points in [-0.6, 0.6] bohr
    along the bond, 4 to 12 electronic and 1 or 4 nuclear functions per
    instance, 4 electrons).  Every instance's integrals are built on the
    device (GM3 in FP64 and BF16); then all SCFs run with DIIS and graphs,
    FP64 and incremental mixed precision, one after another and concurrently
    (neorx.xchf.scf_batch: `workers` host threads, one stream per instance).
    Reports instances/s per way and the largest per-instance |E_mixed -
    E_FP64|."""
    from neorx import xchf

    systems = [synth.config5(k, n) for k in range(n)]
    t0 = time.perf_counter()
    hs = [neorx.System(s) for s in systems]
    xs = [xchf.XCHF(h, s["nelec"], store=("f64", "bf16")) for h, s in zip(hs, systems)]
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    kw = dict(diis=True, graph=True, maxit=maxit, sw_hi=1e-3, sw_lo=1e-6)
    res = {}
    for mode in ("f64", "inc"):
        xchf.scf_batch(xs, workers, mode=mode, **kw)  # graph captures
        t0 = time.perf_counter()
        seq = [x.scf(mode=mode, **kw) for x in xs]
        t_seq = time.perf_counter() - t0
        t0 = time.perf_counter()
        con = xchf.scf_batch(xs, workers, mode=mode, **kw)
        t_con = time.perf_counter() - t0
        same = all(a["comps"]["E_TOTAL"] == b["comps"]["E_TOTAL"] for a, b in zip(seq, con))
        res[mode] = dict(seq=t_seq, con=t_con, energies=[r["comps"]["E_TOTAL"] for r in con],
                         iters=sum(r["niter"] for r in con), same=same)
        print(f"# C5 {mode}: {n} instances sequential {t_seq:.3f} s, concurrent ({workers} threads) {t_con:.3f} s, "
              f"{res[mode]['iters']} iterations", file=sys.stderr, flush=True)
    for x in xs:
        x.close()
    for h in hs:
        h.close()
    de = max(abs(a - b) for a, b in zip(res["inc"]["energies"], res["f64"]["energies"]))
    return {
        "workload": (f"{n} NEO-XCHF RHF instances, proton scanned over [-0.6, 0.6] bohr, nebf "
                     f"{min(s['nebf'] for s in systems)}-{max(s['nebf'] for s in systems)}, npbf "
                     f"{min(s['npbf'] for s in systems)}-{max(s['npbf'] for s in systems)}, 4 electrons (basis sizes "
                     "reduced from config 5's 20-60 / 6-16)"),
        "integral_build_s": t_build,
        "instances_per_s_fp64_sequential": n / res["f64"]["seq"],
        "instances_per_s_fp64_concurrent": n / res["f64"]["con"],
        "instances_per_s_mixed_sequential": n / res["inc"]["seq"],
        "instances_per_s_mixed_concurrent": n / res["inc"]["con"],
        "concurrent_equals_sequential": res["f64"]["same"] and res["inc"]["same"],
        "max_dE_mixed_vs_fp64_Eh": de,
        "energy_range_fp64_Eh": [min(res["f64"]["energies"]), max(res["f64"]["energies"])],
        "workers": workers,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsample", type=int, default=1 << 20, help="e-e-p and e-e-e-p tuples per step")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fock", type=str, default="24,8",
                    help="nebf,npbf of the Fock-build contraction timing (FP64 vs mixed); 'off' to skip")
    ap.add_argument("--scf", type=str, default="c1", choices=("off", "c1"),
                    help="SCF s/iter of the C1 system, FP64 vs mixed precision (builds C1's integrals first)")
    ap.add_argument("--c3", type=str, default="on", choices=("off", "on"),
                    help="config 3: RXCHF-ne contractions at 48/12 functions, FP64 vs FP32 vs BF16 storage")
    ap.add_argument("--n5", type=str, default="on", choices=("off", "on"),
                    help="config 5 at reduced basis sizes: 64 SCF instances, sequential vs concurrent")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device-to-device (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        a = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8}))
        return
    if args.scf != "off":
        os.environ.setdefault("NEORX_VERBOSE", "1")  # the C1 integral build reports its progress (~4 min)
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; multi-GPU runs are not part of it "
              "(the tuple space shards with no communication, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import neorx
    from neorx import synth

    torch.cuda.init()
    arch = neorx.init(0)
    sysd = synth.config2()
    S = neorx.System(sysd)
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream
    ne, npb = sysd["npebf"], sysd["npbf"]
    n1 = ne * ne * npb * npb
    ns = args.nsample
    idx2 = torch.from_numpy(sample_tuples(np, sysd, 2, ns, 2)).cuda()
    idx3 = torch.from_numpy(sample_tuples(np, sysd, 3, ns, 3)).cuda()
    e1 = torch.empty(n1, dtype=torch.float64, device="cuda")
    s1 = torch.empty(n1, dtype=torch.float64, device="cuda")
    e2 = torch.empty(ns, dtype=torch.float64, device="cuda")
    s2 = torch.empty(ns, dtype=torch.float64, device="cuda")
    e3 = torch.empty(ns, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    ev = {c: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for c in (1, 2, 3)}
    per_class_ms = {1: [], 2: [], 3: []}

    def step(timed):
        with torch.cuda.stream(stream):
            if timed:
                ev[1][0].record(stream)
            S.prim_range_dev(1, 0, n1, e1.data_ptr(), s1.data_ptr(), sptr)
            if timed:
                ev[1][1].record(stream)
                ev[2][0].record(stream)
            S.prim_list_dev(2, ns, idx2.data_ptr(), e2.data_ptr(), s2.data_ptr(), sptr)
            if timed:
                ev[2][1].record(stream)
                ev[3][0].record(stream)
            S.prim_list_dev(3, ns, idx3.data_ptr(), e3.data_ptr(), 0, sptr)
            if timed:
                ev[3][1].record(stream)

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(True)
        stream.synchronize()
        for c in (1, 2, 3):
            per_class_ms[c].append(ev[c][0].elapsed_time(ev[c][1]))
        print(f"# step {k + 1}/{args.steps}: " + " ".join(f"cls{c} {per_class_ms[c][-1]:.1f} ms" for c in (1, 2, 3)),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    tuples = n1 + 2 * ns
    value = tuples / (ms_step / 1e3)
    rates = {c: (n1 if c == 1 else ns) / (np.mean(per_class_ms[c]) / 1e3) for c in (1, 2, 3)}
    assert np.isfinite(e2.cpu().numpy()).all() and np.isfinite(e3.cpu().numpy()).all()

    # roofline of the dominant kernel (FP64 vector/matrix peak 78.6 TF/s)
    dom = max((1, 2, 3), key=lambda c: np.mean(per_class_ms[c]))
    flops_per = FLOPS_PER_TUPLE.get(dom)
    roof = None
    if flops_per:
        ntup = n1 if dom == 1 else ns
        ach = flops_per * ntup / (np.mean(per_class_ms[dom]) / 1e3) / 1e12
        roof = {"bound": "fp64_valu", "achieved": round(ach, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 5),
                "traffic": TRAFFIC.get(dom) if args.nsample == (1 << 20) else None,
                "kernel": {1: "k_prim_bin + k_prim_scatter + k_prim_tiers<1,SRC_SORTED>",
                           2: "k_prim_bin + k_prim_scatter + k_prim_tiers<2,SRC_LIST>",
                           3: "k_prim_bin + k_prim_scatter + k_prim_tiers<3,SRC_LIST>"}[dom],
                "flops_per_tuple": flops_per}

    fock = None
    if args.fock != "off":
        fne, fnp = (int(x) for x in args.fock.split(","))
        print(f"# Fock-build contractions nebf={fne} npbf={fnp}", file=sys.stderr, flush=True)
        fock = fock_bench(np, torch, neorx, synth, fne, fnp, 3)

    c3 = None
    if args.c3 == "on":
        print("# C3 RXCHF-ne contractions (FP64 vs FP32 vs BF16 storage)", file=sys.stderr, flush=True)
        c3 = c3_bench(np, torch, neorx, synth)
        print(f"# C3 GM2 built in {c3['gam2_ic1_build_s']:.1f} s; pass {c3['pass_ms_fp64']:.3f} / "
              f"{c3['pass_ms_fp32']:.3f} / {c3['pass_ms_bf16']:.3f} ms", file=sys.stderr, flush=True)

    n5 = None
    if args.n5 == "on":
        print("# C5 batched instances (reduced basis sizes)", file=sys.stderr, flush=True)
        n5 = n5_bench(np, torch, neorx, synth)

    scf = None
    if args.scf == "c1":
        print("# C1 SCF (FP64 vs mixed precision)", file=sys.stderr, flush=True)
        scf = scf_bench(np, torch, neorx, synth)

    cpu = None
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, 16)
        kind, crates, detail = cpu_baseline(np, sysd, args.cpu_budget, cores)
        t_cpu = n1 / crates[1] + ns / crates[2] + ns / crates[3]
        cpu = {"value": tuples / t_cpu, "unit": "integrals/s", "cores": cores, "kind": kind,
               "sample": (f"{detail[1][0]} e-p, {detail[2][0]} e-e-p, {detail[3][0]} e-e-e-p tuples drawn from the "
                          f"same config in {sum(d[1] for d in detail.values()):.1f} s; rates combined in the step's "
                          f"mix"),
               "per_class": {"e-p": crates[1], "e-e-p": crates[2], "e-e-e-p": crates[3]}}

    out = {
        "metric": METRIC, "value": value, "unit": "integrals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "integral kernels: 60 e-prims (s/p/d, 4 centres), 16 p-prims (l<=2), 4-term "
                               "geminal; all e-p tuples + seeded samples of e-e-p and e-e-e-p tuples",
                   "e_p_tuples": n1, "e_e_p_tuples": ns, "e_e_e_p_tuples": ns, "device": arch},
        "per_class_integrals_per_s": {"e-p": rates[1], "e-e-p": rates[2], "e-e-e-p": rates[3]},
        "roofline": roof, "cpu_baseline": cpu, "fock_build": fock, "scf": scf, "c3_rxchfne": c3, "c5_batched": n5,
    }
    print(json.dumps(out))


# Algorithmic FP64 operations per primitive tuple of this workload, counted by
# the host build of the engine over the coefficients the operands' degrees
# make nonzero (tests/count_flops.py); traffic = HBM bytes per launch of the
# dominant kernel from rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/), scaled
# by the copy-kernel calibration.
FLOPS_PER_TUPLE = {1: 56733.7, 2: 113053.3, 3: 33681.3}
TRAFFIC = {2: 366.37e6}  # e-e-p launch group (bin + scatter + tiers) at the default 1,048,576 sampled tuples

if __name__ == "__main__":
    main()
