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


def bucket(n):
    """Block-size bucket of the kernels' template instance for state dimension n."""
    return next(b for b in (8, 16, 32, 48, 64) if n <= b)


def kernel_label(slot, pr, cov=True):
    """The kernel (as rocprofv3 names it) behind a timing slot of
    ukhip_batch_smooth_oddeven_timed for problem pr."""
    n = pr.n[0]
    nb = bucket(n)
    if slot == "k_weigh":
        pk = getattr(pr, "packed", None)
        if pk:
            all_w = pk["K_type"] == "w" and pk["C_type"] == "w"
        else:
            all_w = all(t in (None, "w") for t in pr.K_type) and all(t in (None, "w") for t in pr.C_type)
        return "k_weigh_diag<false>" if all_w else f"k_weigh<{bucket(max(n, max(m_list_of(pr))))}>"
    if slot == "k_backsub":
        if nb >= 32:
            return f"k_backsub_mfma<{nb}>" if cov else f"k_estimate<{nb}>"
        return f"k_backsub<{nb}>"
    # the phase kernels' template arguments: bucket, scalar, workgroups per CU
    # and columns per lane group of the launch variant, and for phases 1 and 3
    # whether the level's observation rows fill at most half a block (level 0:
    # m rows, the levels above: n)
    phase = int(slot[-1])
    if nb >= 32:
        v = device.launch_table()[f"phase{phase}<{nb}, double>"]
        cpl, wg = int(v[1]), int(v[3])
    else:
        cpl, wg = 1, 0
    args = f"{nb}, double, {wg}, {cpl}"
    if phase != 2:  # the level's C blocks: 1 at most half a block of rows (level 0), 2 triangular (above)
        args += f", {1 if 2 * max(m_list_of(pr)) <= nb else 0}|2"
    else:  # DIAG: level 0's D blocks are diagonal (every H the identity, every K 'w')
        pk = getattr(pr, "packed", None)
        if pk:
            lv0 = pk.get("H") is None and pk["K_type"] == "w"
        else:
            lv0 = (all(h is None or np.array_equal(h, np.eye(n)) for h in pr.H)
                   and all(t in (None, "w") for t in pr.K_type))
        args += ", true|false" if lv0 else ", false"
    return f"{slot}<{args}>"


def base_label(label):
    """A kernel label without its launch-variant arguments (the key of the
    traffic summary)."""
    import re
    m = re.match(r"^(k_phase\d<\d+, (?:double|float))(?:, [^>]*)?>$", label)
    return m.group(1) + ">" if m else label


def traffic_of(kernel, workload):
    """HBM bytes per step of `kernel` on `workload` from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE runs, calibrated with a copy kernel),
    or None when no summary covers it."""
    if workload is None:
        return None
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get(workload, {}).get(kernel)
    if not e:
        return None
    return {"bytes_per_step": e["bytes_per_step"], "fetch_bytes": e["fetch_bytes"], "write_bytes": e["write_bytes"],
            "algorithmic_bytes_per_step": e.get("algorithmic_bytes_per_step"), "source": "profiles/traffic.json"}


def copy_calibration(mib):
    """One device-to-device copy (the HIP runtime's copy kernel) of `mib` MiB
    whose source was last touched before a 1 GiB fill, so that it is read
    from HBM and not from a cache."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    size = mib * 1048576
    bufs = [C.c_void_p() for _ in range(3)]
    sizes = (size, 1 << 30, size)
    for p, sz in zip(bufs, sizes):
        if hip.hipMalloc(C.byref(p), sz) != 0:
            raise RuntimeError("hipMalloc failed in the copy calibration")
    hip.hipMemset(bufs[0], 1, size)
    hip.hipMemset(bufs[1], 2, 1 << 30)
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(bufs[2], bufs[0], size, 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed in the copy calibration")
    hip.hipDeviceSynchronize()
    for p in bufs:
        hip.hipFree(p)
    print(f"[bench] copy calibration: {mib} MiB = {size} bytes copied", file=sys.stderr, flush=True)


def m_list_of(pr):
    return [0 if o is None else o.shape[0] for o in pr.o]


def cpu_baseline(name, sample_k, algorithm="oddeven"):
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
    m0 = 0 if pr.o[0] is None else pr.o[0].shape[0]
    if algorithm == "associative" and m0 < pr.n[0]:
        ref = None  # the reference's associative prior needs m_0 >= n (it asserts otherwise)
    if ref is None:
        lib = uk_libs.oracle()
        lib.lib.oracle_set_threads(1)
        kind = "port"
    eqs = capi.EquationSet(pr)
    fn = lib.oe if algorithm == "oddeven" else lib.asc
    t0 = time.perf_counter()
    fn(0, eqs.ptrs, pr.k)
    dt = time.perf_counter() - t0
    eqs.collect(lib.free)
    why = " (a power of two: the reference odd-even aborts on other lengths)" if algorithm == "oddeven" else ""
    if kind == "port" and algorithm == "associative" and m0 < pr.n[0]:
        why += f"; the oracle's restatement, since the reference's associative smoother asserts when m_0={m0} < n"
    return {"value": pr.k / dt, "unit": "timesteps/s", "cores": 1, "kind": kind,
            "sample": f"{WORKLOADS[name]['desc'].split(' k=')[0]} k={sample_k}{why}, "
                      f"kalman_smooth_{algorithm} with covariances, {dt:.1f} s"}


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


def step_rel_err(x, ref):
    """Largest per-step normwise relative error, max_i |x_i - ref_i| / |ref_i|
    (max norms); inf when x holds a non-finite value."""
    if not np.all(np.isfinite(x)):
        return float("inf")
    d = np.max(np.abs(ref), axis=1)
    e = np.max(np.abs(x - ref), axis=1)
    return float(np.max(np.where(d > 0, e / np.where(d > 0, d, 1.0), e)))


def _time(fn, steps, warmup):
    for i in range(warmup):
        fn()
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        fn()
    device.lib().ukhip_synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def mixed_precision_leg(b, timesteps, x64, args):
    """FP32 elimination + FP64 refinement against plain FP64 odd-even on the
    same resident batch: estimates only (the mixed path computes no
    covariances, so its FP64 counterpart is KALMAN_NO_COVARIANCE). Errors are
    per-step normwise relative errors of the estimates against the FP64
    smoother's (x64; lists of per-problem arrays)."""
    from ukhip import capi
    ms64 = _time(lambda: b.smooth("oddeven", options=capi.KALMAN_NO_COVARIANCE, sync=False), args.steps, 1)
    out = {"fp64_states_only": {"ms_per_step": ms64, "value": timesteps / (ms64 * 1e-3)}}
    for r in args.refinements:
        b.smooth_mixed(r)
        err = max(step_rel_err(b.download(i, want_cov=False)[0], x64[i]) for i in range(len(x64)))
        ms = _time(lambda: b.smooth_mixed(r, sync=False), args.steps, 1)
        out[f"refine{r}"] = {"ms_per_step": ms, "value": timesteps / (ms * 1e-3),
                             "speedup_vs_fp64_states_only": ms64 / ms, "max_rel_err_states": err}
        print(f"[bench] mixed, {r} refinement(s): {ms:.2f} ms/step ({ms64 / ms:.2f}x FP64 states-only), "
              f"max rel err {err:.2e}", file=sys.stderr, flush=True)
    return out


def precision_workload(args):
    """The ill-conditioned precision workload (n=64, m=32, k=200000, F = Q
    diag(logspace(0,-3,64)), K = 1e-2 I, C = I): FP64 with covariances, FP64
    estimates only, and FP32 + FP64 refinement."""
    pr = make_problem("n64m32_k200000", args.k)
    b = device.Batch([pr])
    ms_cov = _time(lambda: b.smooth("oddeven", sync=False), args.steps, 1)
    x64 = [b.download(0, want_cov=False)[0]]
    res = {"workload": WORKLOADS["n64m32_k200000"]["desc"] + ("" if args.k is None else f" (k = {pr.k})"),
           "fp64_with_covariances": {"ms_per_step": ms_cov, "value": pr.k / (ms_cov * 1e-3)}}
    res.update(mixed_precision_leg(b, pr.k, x64, args))
    b.free()
    return res


def run_serving(args):
    """The serving workload on a bounded sample: --serve-requests requests
    (n uniform over {3..16}, m uniform over [1,n], lengths log-uniform in
    [100, 50000], F = 0.99 Q_n, G ~ N(0,1/n), K = I, C = 0.1 I, 5% of steps
    unobserved), drawn in turn from a pool of --serve-pool generated requests.
    For each mode (FP64 with covariances; FP32 + refinement, estimates) the
    stream is first submitted at once (capacity), then as a Poisson stream at
    --serve-load times the FP64 capacity (requests per second; the same
    arrival times for both modes). Latency is the server's own, from submit to
    results ready on the host."""
    import threading
    specs = P.serve_specs(args.serve_requests, seed=5)
    t = time.perf_counter()
    pool = [P.serve_problem(s, lists=False) for s in specs[:args.serve_pool]]
    reqs = [pool[i % len(pool)] for i in range(args.serve_requests)]
    steps = sum(p.k for p in reqs)
    print(f"[bench] serving: {len(reqs)} requests ({steps} steps) from a pool of {len(pool)}, generated in "
          f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)

    def stream(refine, arrivals):
        srv = device.Server(hbm_budget_bytes=int(args.serve_budget_gb * 2**30), max_batch=args.serve_max_batch,
                            max_inflight=args.serve_inflight, mixed_refinements=refine)
        lat = [0.0] * len(reqs)
        ids = [None] * len(reqs)
        ready = [threading.Event() for _ in reqs]

        def submit():
            t0 = time.perf_counter()
            for i, p in enumerate(reqs):
                if arrivals is not None:
                    d = t0 + arrivals[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                ids[i] = srv.submit(p)
                ready[i].set()

        t0 = time.perf_counter()
        th = threading.Thread(target=submit)
        th.start()
        for i in range(len(reqs)):
            ready[i].wait()
            lat[i] = srv.wait(ids[i], want_cov=False, want_states=False)[2]
        wall = time.perf_counter() - t0
        th.join()
        srv.close()
        lat = np.array(lat)
        return {"timesteps_per_s": steps / wall, "requests_per_s": len(reqs) / wall, "wall_s": wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "max_ms": float(lat.max())}

    out = {"requests": len(reqs), "pool": len(pool), "timesteps": steps, "hbm_budget_gb": args.serve_budget_gb,
           "max_batch": args.serve_max_batch, "max_inflight": args.serve_inflight}
    modes = (("fp64", -1), (f"mixed_refine{args.serve_refine}", args.serve_refine))
    for name, refine in modes:
        out[name] = {"all_at_once": stream(refine, None)}
        print(f"[bench] serving {name}, all at once: {out[name]['all_at_once']}", file=sys.stderr, flush=True)
    lam = args.serve_load * out["fp64"]["all_at_once"]["requests_per_s"]
    gaps = P.rng_for(77).exponential(1.0 / lam, len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    out["poisson_rate_requests_per_s"] = lam
    out["offered_timesteps_per_s"] = lam * steps / len(reqs)
    for name, refine in modes:
        out[name]["poisson"] = stream(refine, arrivals)
        print(f"[bench] serving {name}, Poisson: {out[name]['poisson']}", file=sys.stderr, flush=True)
    return out


def run_filter(args):
    """The batched sequential filter: --filter-count UltimateKalman filters
    (n = --filter-n, m = n/2, F = 0.98 Q per filter, G ~ N(0,1/n) per step,
    K = I, C = 0.1 I as 'w' inverse standard deviations) advanced
    --filter-steps steps; each step uploads every filter's inputs from host
    memory (PCIe included) and runs one launch."""
    nf, n = args.filter_count, args.filter_n
    m = max(1, n // 2)
    rng = P.rng_for(4242)
    F = np.stack([0.98 * P.random_orthogonal(P.rng_for(5000 + p), n) for p in range(min(nf, 64))])
    F = np.concatenate([F] * ((nf + len(F) - 1) // len(F)))[:nf]
    K = np.ones((nf, n))
    Cw = np.full((nf, m), 1 / np.sqrt(0.1))
    steps = [dict(G=rng.standard_normal((nf, m, n)) / np.sqrt(n), o=rng.standard_normal((nf, m)))
             for _ in range(8)]
    flt = device.Filter(nf, n, m)
    for i in range(2):
        flt.step(F=F, K=K, C=Cw, **steps[i % 8])
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for i in range(args.filter_steps):
        flt.step(F=F, K=K, C=Cw, **steps[i % 8])
    x, _ = flt.estimates(want_R=False)
    wall = time.perf_counter() - t0
    flt.close()
    assert np.all(np.isfinite(x))
    print(f"[bench] filter: {nf} filters n={n} m={m}, {args.filter_steps} steps in {wall:.3f} s", file=sys.stderr,
          flush=True)
    return {"filters": nf, "n": n, "m": m, "steps": args.filter_steps, "ms_per_step": wall * 1e3 / args.filter_steps,
            "value": nf * args.filter_steps / wall, "unit": "filter-steps/s",
            "note": "host inputs uploaded every step (PCIe included)"}


def device_bytes_per_step(n, m):
    """libukhip's odd-even device memory per step (inputs, two chains, factors,
    cross covariances, outputs), in bytes."""
    return 8 * (11.5 * n * n + 9 * n + 2 * m)


def run_batch(args):
    """The batching workload: --batch-count problems, n uniform over
    {4,...,32}, m = n/2, lengths log-uniform in [64, 16384], F = s Q, G ~ N(0,1/n),
    K = C = I. Problems are cut, in order, into chunks that fit the HBM budget;
    each chunk is uploaded, smoothed once untimed (graph capture), then smoothed
    --steps times with its inputs resident; a step of this workload is one
    smoothing of every problem, its time the sum over chunks."""
    if args.launch_table:
        device.set_launch_table(json.loads(args.launch_table))
    elif not args.no_autotune:
        t = time.perf_counter()
        table = device.autotune(32, 65536)
        print(f"[bench] autotuned the QR launch table in {time.perf_counter() - t:.1f}s: {table}", file=sys.stderr,
              flush=True)
    specs = P.batch_specs(count=args.batch_count)
    budget = args.hbm_budget_gb * 1e9
    chunks, cur, used = [], [], 0.0
    for sp in specs:
        need = sp["k"] * device_bytes_per_step(sp["n"], sp["m"])
        if cur and used + need > budget:
            chunks.append(cur)
            cur, used = [], 0.0
        cur.append(sp)
        used += need
    if cur:
        chunks.append(cur)
    total_steps = sum(sp["k"] for sp in specs)
    ms_total, flops_total = 0.0, 0.0
    per_chunk = []
    mixed_acc = {}
    for ci, ch in enumerate(chunks):
        t = time.perf_counter()
        probs = [P.batch_problem(sp, lists=False) for sp in ch]
        tg = time.perf_counter() - t
        t = time.perf_counter()
        b = device.Batch(probs)
        tu = time.perf_counter() - t
        for i in range(max(args.warmup, 1)):
            b.smooth("oddeven")
        device.lib().ukhip_synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            b.smooth("oddeven", sync=False)
        device.lib().ukhip_synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        ksum = sum(sp["k"] for sp in ch)
        if not args.no_mixed:
            x64 = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
            leg = mixed_precision_leg(b, ksum, x64, args)
            for key, v in leg.items():
                acc = mixed_acc.setdefault(key, {"ms_per_step": 0.0, "max_rel_err_states": 0.0})
                acc["ms_per_step"] += v["ms_per_step"]
                acc["max_rel_err_states"] = max(acc["max_rel_err_states"], v.get("max_rel_err_states", 0.0))
        b.free()
        for sp in ch:
            flops_total += _scaled_flops(sp)
        ms_total += ms
        per_chunk.append({"problems": len(ch), "timesteps": ksum, "ms": ms})
        print(f"[bench] chunk {ci + 1}/{len(chunks)}: {len(ch)} problems, {ksum} steps, generated {tg:.1f}s, "
              f"uploaded {tu:.1f}s, {ms:.2f} ms per smoothing", file=sys.stderr, flush=True)
    value = total_steps / (ms_total * 1e-3)
    for key, acc in mixed_acc.items():
        acc["value"] = total_steps / (acc["ms_per_step"] * 1e-3)
        if key != "fp64_states_only":
            acc["speedup_vs_fp64_states_only"] = mixed_acc["fp64_states_only"]["ms_per_step"] / acc["ms_per_step"]
        else:
            del acc["max_rel_err_states"]
    return {
        "metric": METRIC, "value": value, "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_total, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded Philox inputs, i.i.d. N(0,1) observations)",
        "config": {"workload": f"{args.batch_count} independent problems, n uniform over {{4,...,32}}, m = n/2, "
                               f"lengths log-uniform in [64, 16384], F = s Q (s ~ U[0.95,1]), G ~ N(0,1/n), K = C = I",
                   "algorithm": "odd-even, FP64, with covariances; one launch per level per state dimension, "
                                "dimensions concurrent, replayed as a HIP graph",
                   "problems": args.batch_count, "timesteps": total_steps, "chunks": per_chunk,
                   "hbm_budget_gb": args.hbm_budget_gb},
        "whole_path_tflops": flops_total / (ms_total * 1e-3) / 1e12,
        "mixed_precision": mixed_acc or None,
        "qr_launch_table": device.launch_table(),
        "qr_launch_table_source": ("set by --launch-table" if args.launch_table else
                                   "built-in" if args.no_autotune else "autotuned (ukhip_autotune)"),
    }


_PER_STEP = {}


def _scaled_flops(sp):
    """Algorithmic FLOPs of one problem: the per-step count of a 4096-step
    replay times its length (the per-step count is constant away from the
    few top levels)."""
    key = (sp["n"], sp["m"])
    if key not in _PER_STEP:
        k = 4096
        _PER_STEP[key] = flops.oddeven_flops(sp["n"], [sp["m"]] * k, ["w"] * k, ["w"] * k, cov=True)["total"] / k
    return _PER_STEP[key] * sp["k"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="n48m24_k200000", choices=sorted(WORKLOADS) + ["batch4096"])
    ap.add_argument("--batch-count", type=int, default=4096, help="batch4096: number of problems")
    ap.add_argument("--hbm-budget-gb", type=float, default=100.0,
                    help="batch4096: device memory per resident chunk of problems")
    ap.add_argument("--k", type=int, default=None, help="override the number of timesteps (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-k", type=int, default=8192)
    ap.add_argument("--no-associative", action="store_true", help="skip the associative smoother's timing")
    ap.add_argument("--assoc-steps", type=int, default=2)
    ap.add_argument("--cpu-sample-k-assoc", type=int, default=2048)
    ap.add_argument("--refinements", type=lambda s: [int(v) for v in s.split(",")], default=[1, 2],
                    help="mixed-precision legs: FP64 refinement steps after the FP32 elimination")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--no-precision", action="store_true", help="skip the n=64 precision workload")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-filter", action="store_true", help="skip the batched sequential filter")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in QR launch table")
    ap.add_argument("--launch-table", default=None, metavar="JSON",
                    help='set QR launch-table entries instead of autotuning, e.g. \'{"phase2<48, double>": "c2w2"}\' '
                         '(c: columns per lane group, w: workgroups per CU)')
    ap.add_argument("--filter-count", type=int, default=4096)
    ap.add_argument("--filter-n", type=int, default=16)
    ap.add_argument("--filter-steps", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-pool", type=int, default=250)
    ap.add_argument("--serve-budget-gb", type=float, default=32.0)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-refine", type=int, default=2)
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="Poisson rate as a fraction of the FP64 server's all-at-once request rate")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the timed region, copy MIB MiB device-to-device once (a plain copy kernel of "
                         "known size for calibrating rocprofv3's FETCH_SIZE and WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box here has one MI355X); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch4096":
        print(json.dumps(run_batch(args)), flush=True)
        return

    t = time.perf_counter()
    table = None
    if args.launch_table:
        device.set_launch_table(json.loads(args.launch_table))
    elif not args.no_autotune:
        n_tune = WORKLOADS[args.workload]["n"] if args.no_precision else max(WORKLOADS[args.workload]["n"], 64)
        table = device.autotune(n_tune, 65536)
        print(f"[bench] autotuned the QR launch table in {time.perf_counter() - t:.1f}s: {table}", file=sys.stderr,
              flush=True)
    t = time.perf_counter()
    pr = make_problem(args.workload, args.k)
    print(f"[bench] generated {pr.name} k={pr.k} in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    t = time.perf_counter()
    b = device.Batch([pr])
    print(f"[bench] uploaded in {time.perf_counter() - t:.1f}s ({b.levels} levels)", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        b.smooth("oddeven")
    print("[bench] warmup done", file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    # Timed region: K complete odd-even smoothings, inputs resident in HBM. Each
    # is enqueued with HIP events around every launch (library stream), so the
    # same K runs give the wall-clock figure and the per-kernel durations.
    kern = {name: 0.0 for name in device.OE_KERNELS}
    launches = None
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ms, launches = b.smooth_timed(0)
        for name in kern:
            kern[name] += ms[name] / args.steps
    device.lib().ukhip_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    print(f"[bench] {ms_per_step:.2f} ms/step; events: " +
          ", ".join(f"{k} {v:.3f}" for k, v in kern.items()), file=sys.stderr, flush=True)

    n, ml = pr.n[0], m_list_of(pr)
    fl = flops.oddeven_flops(n, ml, pr.K_type, pr.C_type, cov=True)
    fl_of = {"k_weigh": fl["weigh"], "k_phase1": fl["phase1"], "k_phase2": fl["phase2"],
             "k_phase3": fl["phase3"], "k_backsub": fl["backsub"]}
    dom = max(fl_of, key=lambda k: kern[k])
    label = kernel_label(dom, pr)
    achieved = fl_of[dom] / (kern[dom] * 1e-3) / 1e12
    traffic = traffic_of(base_label(label), args.workload if args.k is None else None)
    if traffic is not None and dom == "k_backsub":
        traffic["algorithmic_bytes_per_step"] = flops.oddeven_backsub_bytes(n, pr.k)

    mixed = None
    if not args.no_mixed:
        x64 = [b.download(0, want_cov=False)[0]]
        mixed = mixed_precision_leg(b, pr.k, x64, args)

    assoc = None
    if not args.no_associative:
        for i in range(max(1, args.warmup // 2)):
            b.smooth("associative")
        device.lib().ukhip_synchronize()
        ta = time.perf_counter()
        for i in range(args.assoc_steps):
            b.smooth("associative", sync=False)
        device.lib().ukhip_synchronize()
        ams = (time.perf_counter() - ta) * 1e3 / args.assoc_steps
        assoc = {"algorithm": "associative scan (Saerkkae-Garcia-Fernandez), FP64, with covariances",
                 "ms_per_step": ams, "value": pr.k / (ams * 1e-3), "unit": "timesteps/s", "steps": args.assoc_steps}
        if not args.no_cpu_baseline:
            try:
                assoc["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_sample_k_assoc, "associative")
            except Exception as e:  # the GPU figure stands without it
                assoc["cpu_baseline"] = {"error": str(e)}
        print(f"[bench] associative {ams:.2f} ms/step", file=sys.stderr, flush=True)
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
        "roofline": {"bound": "mfma", "kernel": label,
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "compute": ("FP64 matrix cores, v_mfma_f64_16x16x4f64" if "mfma" in label else "FP64 VALU") +
                                " (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "algorithmic_flops_per_step": fl_of[dom], "launches_per_step": launches[dom],
                     "kernel_ms_per_step": kern[dom], "avg_launch_ms": kern[dom] / launches[dom]},
        "kernels_ms_per_step": {k: v for k, v in kern.items()},
        "kernels": {k: kernel_label(k, pr) for k in fl_of},
        "qr_launch_table": table if table is not None else device.launch_table(),
        "qr_launch_table_source": ("autotuned (ukhip_autotune)" if table is not None
                                   else "set by --launch-table" if args.launch_table else "built-in"),
        "kernels_tflops": {k: fl_of[k] / (kern[k] * 1e-3) / 1e12 for k in fl_of if kern[k] > 0},
        "algorithmic_flops_per_timestep": fl["total"] / pr.k,
        "whole_path_tflops": fl["total"] / (kern["total"] * 1e-3) / 1e12,
    }
    if mixed:
        out["mixed_precision"] = mixed
    if assoc:
        out["associative"] = assoc
    if not args.no_mixed and not args.no_precision:
        out["precision_workload"] = precision_workload(args)
    if not args.no_serving:
        out["serving"] = run_serving(args)
    if not args.no_filter:
        out["sequential_filter"] = run_filter(args)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_sample_k)
            out["cpu_port_allcores"] = cpu_port_allcores(args.workload, args.cpu_sample_k)
        except Exception as e:  # the GPU figure stands without it
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
