"""Benchmark of libukhip's odd-even smoother on one MI355X.

Workload (default): one problem with n = 48 states, m = 24 observations per
step, k = 200,000 steps, every 10th step unobserved, F = 0.99 Q (Q random
orthogonal), G_i ~ N(0, 1/48) drawn per step, K = I, C = 0.1 I; FP64;
estimates and covariances. A "step" of this benchmark is one complete
smoothing of that problem with its inputs already resident in HBM.

Prints one JSON line (metric, value = smoothed timesteps/s, roofline of the
dominant kernel class measured with HIP events on the library stream, and
the reference's CPU time on a bounded sample).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ultimate-kalman_amd", "python"))

from ukhip import device, flops  # noqa: E402
from ukhip import problems as P  # noqa: E402

METRIC = "timesteps/s, batched FP32+refine vs plain FP64 odd-even on GPU; p50/p99 latency"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "n48m24_k200000": dict(n=48, m=24, k=200_000, unob10=True, desc="n=48 m=24 k=200000, every 10th step unobserved, "
                           "F=0.99Q, G~N(0,1/48), K=I, C=0.1I"),
    "n6_k100000": dict(n=6, m=6, k=100_000, unob10=False, desc="n=6 m=6 k=100000, F orthogonal, H=G=K=C=I"),
    "n64m32_k200000": dict(n=64, m=32, k=200_000, unob10=False, desc="n=64 m=32 k=200000, F=Q diag(logspace(0,-3)), "
                           "G~N(0,1/64), K=1e-2 I, C=I"),
}


def make_problem(name, k=None):
    w = WORKLOADS[name]
    k = k or w["k"]
    n, m = w["n"], w["m"]
    unob = (lambda i: i % 10 == 9) if w["unob10"] else None
    if n == 48:
        F = 0.99 * P.random_orthogonal(P.rng_for(1002), 48)
        return P.iid_problem(k, 48, 24, seed=13, F=F, G_scale=1 / np.sqrt(48.0), Kvar=1.0, Cvar=0.1,
                             unobserved=unob, name=name)
    if n == 6:
        F = P.random_orthogonal(P.rng_for(1001), 6)
        return P.iid_problem(k, 6, 6, seed=11, F=F, K_kind="w", C_kind="w", name=name)
    F = P.random_orthogonal(P.rng_for(1004), 64) @ np.diag(np.logspace(0, -3, 64))
    return P.iid_problem(k, 64, 32, seed=14, F=F, G_scale=1 / 8.0, Kvar=1e-2, Cvar=1.0, name=name)


def m_list_of(pr):
    return [0 if o is None else o.shape[0] for o in pr.o]


def cpu_baseline(name, sample_k):
    """The reference itself (oracle/_ref: its C smoothers, sequential backend,
    one core) on a bounded sample of the workload; the oracle port if the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import uk_libs
    from ukhip import capi
    pr = make_problem(name, sample_k)
    ref = uk_libs.reference()
    kind = "reference"
    lib = ref
    if ref is None:
        lib = uk_libs.oracle()
        lib.lib.oracle_set_threads(1)
        kind = "port"
    eqs = capi.EquationSet(pr)
    t0 = time.perf_counter()
    lib.oe(0, eqs.ptrs, pr.k)
    dt = time.perf_counter() - t0
    eqs.collect(lib.free)
    return {"value": pr.k / dt, "unit": "timesteps/s", "cores": 1, "kind": kind,
            "sample": f"{WORKLOADS[name]['desc'].split(' k=')[0]} k={sample_k} (a power of two: the reference "
                      f"odd-even aborts on other lengths), kalman_smooth_oddeven with covariances, {dt:.1f} s"}


def cpu_port_allcores(name, sample_k):
    """The oracle's restatement with OpenMP on all host cores (parallel-for over
    the pairs of each level, like the reference's TBB backend)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import uk_libs
    from ukhip import capi
    pr = make_problem(name, sample_k)
    o = uk_libs.oracle()
    cores = min(os.cpu_count() or 1, 64)
    o.lib.oracle_set_threads(cores)
    eqs = capi.EquationSet(pr)
    t0 = time.perf_counter()
    o.oe(0, eqs.ptrs, pr.k)
    dt = time.perf_counter() - t0
    eqs.collect(o.free)
    o.lib.oracle_set_threads(1)
    return {"value": pr.k / dt, "unit": "timesteps/s", "cores": cores, "kind": "port",
            "sample": f"k={sample_k}, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="n48m24_k200000", choices=sorted(WORKLOADS))
    ap.add_argument("--k", type=int, default=None, help="override the number of timesteps (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-k", type=int, default=8192)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box here has one MI355X); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    pr = make_problem(args.workload, args.k)
    print(f"[bench] generated {pr.name} k={pr.k} in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    t = time.perf_counter()
    b = device.Batch([pr])
    print(f"[bench] uploaded in {time.perf_counter() - t:.1f}s ({b.levels} levels)", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        b.smooth("oddeven")
    print("[bench] warmup done", file=sys.stderr, flush=True)

    phase = np.zeros(4)
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        phase += np.array(b.smooth_timed(0))  # HIP events on the library stream around each kernel class
    device.lib().ukhip_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    ms_weigh, ms_elim, ms_back, ms_total = phase / args.steps
    print(f"[bench] {ms_per_step:.2f} ms/step (events: weigh {ms_weigh:.2f}, eliminate {ms_elim:.2f}, "
          f"back {ms_back:.2f}, total {ms_total:.2f})", file=sys.stderr, flush=True)

    n, ml = pr.n[0], m_list_of(pr)
    fl = flops.oddeven_flops(n, ml, pr.K_type, pr.C_type, cov=True)
    cls = "eliminate" if ms_elim >= ms_back else "backsub"
    cls_ms = ms_elim if cls == "eliminate" else ms_back
    achieved = fl[cls] / (cls_ms * 1e-3) / 1e12
    b.free()

    out = {
        "metric": METRIC,
        "value": pr.k / (ms_per_step * 1e-3),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox inputs, i.i.d. N(0,1) observations)",
        "config": {"workload": WORKLOADS[args.workload]["desc"] if args.k is None else
                   f"{WORKLOADS[args.workload]['desc']} (k overridden to {pr.k})",
                   "algorithm": "odd-even (Gargir-Toledo), FP64, with covariances", "problems": 1,
                   "timesteps": pr.k},
        "roofline": {"bound": "mfma", "kernel": f"k_{cls} (all levels)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "algorithmic_flops": fl[cls], "kernel_ms": cls_ms},
        "phases_ms": {"weigh": ms_weigh, "eliminate": ms_elim, "backsub": ms_back, "events_total": ms_total},
        "algorithmic_flops_per_timestep": fl["total"] / pr.k,
        "whole_path_tflops": fl["total"] / (ms_total * 1e-3) / 1e12,
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_sample_k)
            out["cpu_port_allcores"] = cpu_port_allcores(args.workload, args.cpu_sample_k)
        except Exception as e:  # the GPU figure stands without it
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
