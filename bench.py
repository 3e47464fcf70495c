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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsample", type=int, default=1 << 20, help="e-e-p and e-e-e-p tuples per step")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline timing")
    ap.add_argument("--no-cpu", action="store_true")
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
        roof = {"bound": "mfma", "achieved": round(ach, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 5),
                "traffic": TRAFFIC.get(dom) if args.nsample == (1 << 20) else None,
                "kernel": {1: "k_prim_small<1,SRC_SORTED>", 2: "k_prim_small<2,SRC_LIST> + k_prim_gen<2,SRC_LIST>",
                           3: "k_prim_small<3,SRC_LIST> + k_prim_gen<3,SRC_LIST>"}[dom],
                "flops_per_tuple": flops_per}

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
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


# Algorithmic FP64 operations per primitive tuple of this workload, counted by
# the host build of the engine over the coefficients the operands' degrees
# make nonzero (tests/count_flops.py); traffic = HBM bytes per launch of the
# dominant kernel from rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/), scaled
# by the copy-kernel calibration.
FLOPS_PER_TUPLE = {1: 99762.1, 2: 251422.5, 3: 46648.4}
TRAFFIC = {2: 131.9e6}  # e-e-p launch group at the default 1,048,576 sampled tuples

if __name__ == "__main__":
    main()
