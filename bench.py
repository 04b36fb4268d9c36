"""Benchmark of libukhip on one MI355X: batched FP32 elimination + FP64
refinement against plain FP64 odd-even smoothing, on the batching workload.

Headline (one JSON line on stdout): timesteps/s of the batched mixed-precision
odd-even smoother (FP32 elimination, FP64 iterative refinement of the
estimates) over 4,096 independent problems (n uniform over {4,8,...,32},
m = n/2, lengths log-uniform in [64, 16384], F = s Q with s ~ U[0.95, 1],
G ~ N(0, 1/n) per step, K = C = I), with its inputs resident in HBM. A "step"
of the timed region is one smoothing of every problem. Beside it, on the same
resident workload: the batched FP64 odd-even smoother (estimates only, and
with covariances), and the plain FP64 odd-even smoother run one problem per
smoothing; the ratios; the estimates' error against the FP64 results; the
roofline of the headline path's dominant kernel from HIP events.

Secondary workloads, each a JSON object in the same line: the single long
problem (n=48, m=24, k=200,000) in FP64 with the roofline of its dominant
kernel, mixed precision, the end-to-end kalman_smooth_oddeven call through the
C ABI, the associative smoother, and the errors against the reference's own
output; the ill-conditioned precision workload (n=64, m=32, k=200,000); a
bounded sample of the serving workload; the batched sequential filter; and the
reference's CPU time on bounded samples.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ultimate-kalman_amd", "python"))

from ukhip import device, flops  # noqa: E402
from ukhip import problems as P  # noqa: E402

METRIC = "timesteps/s, batched FP32+refine vs plain FP64 odd-even on GPU; p50/p99 latency"
# MI355X published peaks (AMD): FP64 vector/matrix, FP32 vector/matrix, HBM3E
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
NOCOV = 1 << 16  # KALMAN_NO_COVARIANCE
COVDIAG = 1 << 25  # UKHIP_COVARIANCE_DIAG: the first level's back substitution forms only diag(S_j)

WORKLOADS = {
    "n48m24_k200000": dict(n=48, m=24, k=200_000, unob10=True, desc="n=48 m=24 k=200000, every 10th step unobserved, "
                           "F=0.99Q, G~N(0,1/48), K=I, C=0.1I"),
    "n6_k100000": dict(n=6, m=6, k=100_000, unob10=False, desc="n=6 m=6 k=100000, F orthogonal, H=G=K=C=I"),
    "n64m32_k200000": dict(n=64, m=32, k=200_000, unob10=False, desc="n=64 m=32 k=200000, F=Q diag(logspace(0,-3)), "
                           "G~N(0,1/64), K=1e-2 I, C=I"),
}
BATCH_DESC = ("4096 independent problems, n uniform over {4,8,...,32}, m = n/2, lengths log-uniform in [64, 16384], "
              "F = s Q (s ~ U[0.95,1]), G ~ N(0,1/n) per step, K = C = I")


def say(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_problem(name, k=None, lists=True):
    w = WORKLOADS[name]
    k = k or w["k"]
    n = w["n"]
    unob = (lambda i: i % 10 == 9) if w["unob10"] else None
    if n == 48:
        F = 0.99 * P.random_orthogonal(P.rng_for(1002), 48)
        return P.iid_problem(k, 48, 24, seed=13, F=F, G_scale=1 / np.sqrt(48.0), Kvar=1.0, Cvar=0.1,
                             unobserved=unob, name=name, lists=lists)
    if n == 6:
        F = P.random_orthogonal(P.rng_for(1001), 6)
        return P.iid_problem(k, 6, 6, seed=11, F=F, K_kind="w", C_kind="w", name=name, lists=lists)
    F = P.random_orthogonal(P.rng_for(1004), 64) @ np.diag(np.logspace(0, -3, 64))
    return P.iid_problem(k, 64, 32, seed=14, F=F, G_scale=1 / 8.0, Kvar=1e-2, Cvar=1.0, name=name, lists=lists)


def bucket(n):
    """Block-size bucket of the kernels' template instance for state dimension n."""
    return next(b for b in (8, 16, 32, 48, 64) if n <= b)


def bucket_qr(n):
    """The bucket of the elimination's and the refinement sweeps' kernels (24 for 17..24)."""
    return next(b for b in (8, 16, 24, 32, 48, 64) if n <= b)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def m_list_of(pr):
    if pr.packed and not pr.o:
        pk = pr.packed
        return [pk["m"] if ob else 0 for ob in pk["observed"]]
    return [0 if o is None else o.shape[0] for o in pr.o]


def kernel_label(slot, n, m, fp32=False, cov=True, all_w=True, h_ident=True, level=None):
    """The kernel (as rocprofv3 names it) behind a timing slot for a part of
    state dimension n with m observation rows; level None covers every level
    (the phase kernels' per-level structure flags shown as *)."""
    nb = bucket(n)
    nq = bucket_qr(n)
    T = "float" if fp32 else "double"
    if slot == "k_weigh":
        return "k_weigh_diag<" + ("true" if fp32 else "false") + ">" if all_w else f"k_weigh<{bucket(max(n, m))}>"
    if slot == "k_backsub":
        if nb >= 32:
            return f"k_backsub_mfma<{nb}>" if cov else f"k_estimate<{nb}>"
        return f"k_backsub<{nb}>"
    if slot == "k_solve_est":
        return f"k_solve_est<{nq}, {T}, *>"
    if slot == "k_normal":
        nm = bucket_qr(max(n, m))
        return f"{'k_normal_w' if nm <= 32 else 'k_normal'}<{nm}, *> + k_gather + k_add"
    if slot == "k_fwd":
        return f"k_fwd_even<{nq}, {T}> + k_fwd_odd<{nq}, {T}>"
    if slot == "convert":
        return "k_to32 / k_to64"
    phase = int(slot[-1])
    if nq >= 24:
        v = device.launch_table()[f"phase{phase}<{nq}, {T}>"]
        cpl, wg = int(v[1]), int(v[3])
    else:
        cpl, wg = 1, 0
    if cpl == 0:
        return f"k_phase2_wy<{nq}, {T}, *>"
    return f"{slot}<{nq}, {T}, {wg}, {cpl}, *>"


def traffic_of(kernel, workload):
    """HBM bytes per step of `kernel` on `workload` from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE runs, calibrated with a copy kernel),
    or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if workload is None or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    e = t.get(workload, {}).get(kernel)
    if not e:
        return None
    return {"bytes_per_step": e["bytes_per_step"], "fetch_bytes": e["fetch_bytes"], "write_bytes": e["write_bytes"],
            "algorithmic_bytes_per_step": e.get("algorithmic_bytes_per_step"), "source": "profiles/traffic.json"}


def headline_traffic(kernel, specs, sample=600):
    """HBM bytes per smoothing of an FP32 elimination kernel over the whole
    batching workload: the committed PMC summary of one eager mixed smoothing
    of the workload's first `sample` problems (bench.py --mixed-chunk-only),
    scaled by the time steps of the kernel's bucket in the workload over those
    in the sample. None when no summary covers the kernel."""
    base = kernel.split(",")[0] + ", " + kernel.split(",")[1].strip() + ">" if kernel.count(",") >= 2 else kernel
    t = traffic_of(base, f"batch_mixed_chunk{sample}")
    if t is None:
        return None
    nb = int(base.split("<")[1].split(",")[0])
    k_sample = sum(sp["k"] for sp in P.batch_specs(count=sample) if bucket_qr(sp["n"]) == nb)
    k_all = sum(sp["k"] for sp in specs if bucket_qr(sp["n"]) == nb)
    if k_sample == 0:
        return None
    f = k_all / k_sample
    alg = None
    if base.startswith("k_phase2<") and ", float>" in base:
        # the FP32 levels (0-7) of the mixed path; level 0 with D diagonal and B formed on load (K, C given as 'w')
        alg = sum(flops.oddeven_phase2_bytes(sp["n"], sp["k"], fp32_levels=8, level0_diag_synth=True)
                  for sp in specs if bucket_qr(sp["n"]) == nb)
    meas = t["bytes_per_step"] * f
    return {"bytes_per_step": meas, "fetch_bytes": t["fetch_bytes"] * f,
            "write_bytes": t["write_bytes"] * f, "algorithmic_bytes_per_step": alg,
            "measured_over_algorithmic": (meas / alg) if alg else None,
            "source": f"profiles/traffic.json, batch_mixed_chunk{sample} ({base}), x{f:.3f}: bucket-{nb} time steps "
                      f"of the workload over the sample's"}


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
    for p, sz in zip(bufs, (size, 1 << 30, size)):
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
    say(f"copy calibration: {mib} MiB = {size} bytes copied")


def step_rel_err(x, ref):
    """Largest per-step normwise relative error, max_i |x_i - ref_i| / |ref_i|
    (max norms, one row per step); inf when x holds a non-finite value."""
    x = np.asarray(x, dtype=np.float64).reshape(len(ref), -1)
    ref = np.asarray(ref, dtype=np.float64).reshape(len(ref), -1)
    if not np.all(np.isfinite(x)):
        return float("inf")
    d = np.max(np.abs(ref), axis=1)
    e = np.max(np.abs(x - ref), axis=1)
    return float(np.max(np.where(d > 0, e / np.where(d > 0, d, 1.0), e)))


def _time(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    device.lib().ukhip_synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


# ---------------------------------------------------------------------------------------------------------------
# algorithmic operation counts
# ---------------------------------------------------------------------------------------------------------------

_PER_STEP = {}


def oddeven_flops_per_step(n, m, cov):
    """Algorithmic FLOPs per step of the odd-even smoother by phase (the LAWN-41
    replay of the reference's call shapes, ukhip/flops.py) for constant n, m
    and 'w' covariances, from a 4096-step replay (the per-step count is
    constant away from the few top levels)."""
    key = (n, m, cov)
    if key not in _PER_STEP:
        k = 4096
        f = flops.oddeven_flops(n, [m] * k, ["w"] * k, ["w"] * k, cov=cov)
        _PER_STEP[key] = {kk: v / k for kk, v in f.items()}
    return _PER_STEP[key]


def refine_bytes_per_step(n, m, slot):
    """Algorithmic HBM bytes per step of one refinement-sweep kernel class: the
    blocks it must read and the vectors it writes (FP64 chain and vectors, FP32
    factors; one even factor slot per step over all levels)."""
    nn = n * n
    if slot == "k_normal":  # reads x (2 steps), B (or F and k), C, c, o; writes gA, gB, g, then x += d
        return 8.0 * (2 * n + nn + m * n + n + m) + 8.0 * 3 * n + 8.0 * 3 * n
    if slot == "k_fwd":  # even: R (FP32), g; odd: Y, B~ (FP32), y; writes y, g'
        return 4.0 * 3 * nn + 8.0 * 4 * n
    if slot == "k_solve_est":  # R, B~, Y (FP32), y, two neighbours' x; writes x
        return 4.0 * 3 * nn + 8.0 * 4 * n
    return None


# ---------------------------------------------------------------------------------------------------------------
# the headline: batching workload
# ---------------------------------------------------------------------------------------------------------------

def device_bytes_per_step(n, m):
    """libukhip's odd-even device memory per step (inputs, two chains, factors,
    cross covariances, outputs, the mixed-precision workspace), in bytes."""
    return 8 * (11.5 * n * n + 9 * n + 2 * m) + 4 * (8.5 * n * n + 6 * n) + 8 * 6 * n


def mixed_bytes_per_step(n, m):
    """libukhip's device memory per step for a batch smoothed only by the
    mixed-precision path for its estimates (raw inputs, the FP64 level-0 chain
    for the residuals, the FP32 chains and factors, the refinement vectors;
    the FP64 elimination's storage is never allocated), in bytes."""
    return 1.05 * (8 * (3 * n * n + m * n + 10.5 * n + 2 * m + 4) + 4 * (7.5 * n * n + 4 * n))


def chunked(specs, bytes_per_step, budget):
    chunks, cur, used = [], [], 0.0
    for sp in specs:
        need = sp["k"] * bytes_per_step(sp["n"], sp["m"])
        if cur and used + need > budget:
            chunks.append(cur)
            cur, used = [], 0.0
        cur.append(sp)
        used += need
    if cur:
        chunks.append(cur)
    return chunks


def accumulate_parts(b, ch, R, kern):
    """Per-kernel device time of one eager mixed smoothing of batch b (chunk
    specs ch; parts one after another, HIP events around every launch),
    added into kern = (ms, launches, flops, bytes) dicts keyed by kernel label.
    Returns the smoothing's summed device time (ms)."""
    slot_ms, slot_launch, slot_flops, slot_bytes = kern
    total = 0.0
    for n_, ms, nl in b.timed_parts(mixed=(R, 0, False)):
        total += ms["total"]
        m_ = max(n_ // 2, 1)
        kpart = sum(sp["k"] for sp in ch if sp["n"] == n_)
        fps = oddeven_flops_per_step(n_, m_, False)
        for slot in device.OE_KERNELS:
            if slot == "total" or ms[slot] <= 0:
                continue
            # the phases of levels 0-7 run in FP32 (the few steps of the levels above in FP64)
            key = kernel_label(slot, n_, m_, fp32=True, cov=False)
            slot_ms[key] = slot_ms.get(key, 0.0) + ms[slot]
            slot_launch[key] = slot_launch.get(key, 0) + nl[slot]
            if slot.startswith("k_phase"):
                slot_flops[key] = slot_flops.get(key, 0.0) + fps["phase" + slot[-1]] * kpart
            rb = refine_bytes_per_step(n_, m_, slot)
            if rb is not None:  # k_solve_est also forms the first estimates: R + 1 sweeps
                slot_bytes[key] = slot_bytes.get(key, 0.0) + rb * kpart * (R + (slot == "k_solve_est"))
    return total


def mixed_pass(specs, args, kern):
    """The headline's timed leg: the mixed-precision path (estimates, R FP64
    refinement steps) over the whole workload in chunks sized for its own
    device memory; returns (ms per smoothing summed over chunks, per-chunk
    records). The per-kernel times of one eager smoothing per chunk go into
    kern (accumulate_parts): the roofline's kernel ran on these chunks. With
    --mixed-parts-only that eager smoothing is the only one, and its summed
    device time stands for the chunk's time (a profiling aid: a kernel trace
    of that run holds exactly the launches the roofline times)."""
    total, rec = 0.0, []
    for ci, ch in enumerate(chunked(specs, mixed_bytes_per_step, args.mixed_hbm_budget_gb * 1e9)):
        t = time.perf_counter()
        probs = [P.batch_problem(sp, lists=False) for sp in ch]
        tg = time.perf_counter() - t
        b = device.Batch(probs)
        if args.mixed_parts_only:
            ms = accumulate_parts(b, ch, args.refine, kern)
        else:
            ms = _time(lambda: b.smooth_mixed_opt(args.refine, 0, False, sync=False), args.steps,
                       max(args.warmup, 1))
            accumulate_parts(b, ch, args.refine, kern)
        b.free()
        del probs
        ksum = sum(sp["k"] for sp in ch)
        total += ms
        rec.append({"problems": len(ch), "timesteps": ksum, "ms_mixed": ms})
        say(f"mixed chunk {ci + 1}: {len(ch)} problems, {ksum} steps, generated {tg:.1f}s; {ms:.1f} ms per smoothing")
    return total, rec


def plain_fp64_pass(probs, steps):
    """The plain FP64 odd-even smoother (estimates only) run one problem per
    smoothing, every problem resident in HBM in a batch of its own: each pass
    smooths the problems one after another; returns ms per pass."""
    L = device.lib()
    bs = [device.Batch([p]) for p in probs]
    try:
        for b in bs:  # capture each problem's graph
            L.ukhip_batch_smooth_oddeven(b.h, NOCOV)
        L.ukhip_synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for b in bs:
                if L.ukhip_batch_smooth_oddeven(b.h, NOCOV) != 0:
                    raise RuntimeError(device.last_error())
        L.ukhip_synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        x0 = bs[0].download(0, want_cov=False)[0]
        return ms, x0
    finally:
        for b in bs:
            b.free()


def run_headline(args):
    specs = P.batch_specs(count=args.batch_count)
    chunks = chunked(specs, device_bytes_per_step, args.hbm_budget_gb * 1e9)
    total_steps = sum(sp["k"] for sp in specs)
    R = args.refine
    slot_ms, slot_launch, slot_flops, slot_bytes = {}, {}, {}, {}
    mixed_ms, mixed_chunks = mixed_pass(specs, args, (slot_ms, slot_launch, slot_flops, slot_bytes))
    acc = {"mixed": 0.0, "fp64_states_only": 0.0, "fp64_with_covariances": 0.0, "plain_fp64_states_only": 0.0,
           "mixed_with_covariances": 0.0}
    err_mixed, err_plain, err_mixcov = 0.0, 0.0, 0.0
    per_chunk = []
    oracle_sample = []
    for ci, ch in enumerate([] if args.headline_mixed_only else chunks):
        t = time.perf_counter()
        probs = [P.batch_problem(sp, lists=False) for sp in ch]
        tg = time.perf_counter() - t
        t = time.perf_counter()
        b = device.Batch(probs)
        tu = time.perf_counter() - t
        ksum = sum(sp["k"] for sp in ch)
        cov_ms = _time(lambda: b.smooth("oddeven", sync=False), args.steps, max(args.warmup, 1))
        est_ms = _time(lambda: b.smooth("oddeven", options=NOCOV, sync=False), args.steps, max(args.warmup, 1))
        x64 = [b.download(i, want_cov=False)[0] for i in range(len(probs))]
        mix_ms = _time(lambda: b.smooth_mixed_opt(R, 0, False, sync=False), args.steps, max(args.warmup, 1))
        err_mixed = max(err_mixed, max(step_rel_err(b.download(i, want_cov=False)[0], x64[i])
                                       for i in range(len(probs))))
        mixcov_ms = _time(lambda: b.smooth_mixed_opt(R, 0, True, sync=False), 1, 1)
        # covariance diagonals of the mixed path against the FP64 ones, on the chunk's first problems
        b.smooth("oddeven")
        d64 = [b.download_cov(i, "D") for i in range(min(8, len(probs)))]
        b.smooth_mixed_opt(R, 0, True)
        err_mixcov = max(err_mixcov, max(step_rel_err(b.download_cov(i, "D"), d64[i]) for i in range(len(d64))))
        if ci == 0:  # batched FP64 against the oracle on a few short problems of the first chunk
            oracle_sample = [(i, probs[i]) for i in sorted(range(len(ch)), key=lambda i: ch[i]["k"])[:4]]
            oracle_sample = [(i, p, x64[i]) for i, p in oracle_sample]
        b.free()
        plain_ms, x0 = plain_fp64_pass(probs, args.steps)
        err_plain = max(err_plain, step_rel_err(x0, x64[0]))
        for key, v in (("mixed", mix_ms), ("fp64_states_only", est_ms), ("fp64_with_covariances", cov_ms),
                       ("plain_fp64_states_only", plain_ms), ("mixed_with_covariances", mixcov_ms)):
            acc[key] += v
        per_chunk.append({"problems": len(ch), "timesteps": ksum, "ms_mixed": mix_ms, "ms_fp64_states_only": est_ms,
                          "ms_plain_fp64_states_only": plain_ms})
        say(f"chunk {ci + 1}/{len(chunks)}: {len(ch)} problems, {ksum} steps, generated {tg:.1f}s, uploaded "
            f"{tu:.1f}s; ms per smoothing: mixed {mix_ms:.1f}, FP64 states {est_ms:.1f}, FP64 cov {cov_ms:.1f}, "
            f"plain FP64 states {plain_ms:.1f}")
    # the oracle on the sampled problems (lists rebuilt from the same seeds)
    oracle_err = None
    if not args.no_oracle_sample and oracle_sample:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import uk_libs
        o = uk_libs.oracle()
        o.lib.oracle_set_threads(min(16, os.cpu_count() or 1))
        oracle_err = 0.0
        for i, _, x in oracle_sample:
            pr = P.batch_problem(chunks[0][i], lists=True)
            ox, _ = o.oddeven(pr, NOCOV)
            oracle_err = max(oracle_err, step_rel_err(x, np.stack(ox)))
        o.lib.oracle_set_threads(1)
    dom = max(slot_ms, key=lambda k: slot_ms[k])
    if dom in slot_flops:
        achieved = slot_flops[dom] / (slot_ms[dom] * 1e-3) / 1e12
        peak, unit, bound = (FP32_PEAK_TFLOPS if "float" in dom else FP64_PEAK_TFLOPS), "TFLOP/s", "valu"
    else:
        achieved = (slot_bytes.get(dom, 0.0)) / (slot_ms[dom] * 1e-3) / 1e9
        peak, unit, bound = HBM_PEAK_GBS, "GB/s", "hbm"
    acc["mixed_at_fp64_chunking"] = acc.pop("mixed")
    acc["mixed"] = mixed_ms
    value = total_steps / (acc["mixed"] * 1e-3)
    rates = {k: (total_steps / (v * 1e-3) if v > 0 else None) for k, v in acc.items()}
    ratio = (lambda a, b: rates[a] / rates[b] if rates[a] and rates[b] else None)
    roof = {"bound": bound, "kernel": dom, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": headline_traffic(dom, specs) if dom in slot_flops else None,
            "launches_per_step": slot_launch[dom], "kernel_ms_per_step": slot_ms[dom],
            "avg_launch_ms": slot_ms[dom] / max(slot_launch[dom], 1),
            "algorithmic_flops_per_step": slot_flops.get(dom), "algorithmic_bytes_per_step": slot_bytes.get(dom),
            "measured": "HIP events around every launch of one eager mixed smoothing of each of the timed "
                        "region's chunks, parts one after another (the timed region replays graphs with the parts "
                        "on concurrent streams); the launches are the FP32 levels' and, for the problems with more "
                        "than 8 levels, the FP64 top levels' (a few steps each), as the flops count",
            "note": ("FP32 peak is the published 157.3 TFLOP/s, reached only with packed FP32, which this build "
                     "disables; the unpacked FP32 VALU rate equals the FP64 one, 78.6") if "float" in dom else ""}
    kern = {k: {"ms_per_step": v, "launches": slot_launch[k],
                "tflops": (slot_flops[k] / (v * 1e-3) / 1e12) if k in slot_flops else None}
            for k, v in sorted(slot_ms.items(), key=lambda kv: -kv[1])}
    return {
        "value": value, "ms_per_step": acc["mixed"], "total_steps": total_steps, "roofline": roof,
        "headline": {
            "refinements": R,
            "ratio_vs_plain_fp64": ratio("mixed", "plain_fp64_states_only"),
            "ratio_vs_batched_fp64": ratio("mixed", "fp64_states_only"),
            "timesteps_per_s": rates,
            "ms_per_step": acc,
            "max_rel_err_states_vs_fp64": err_mixed,
            "max_rel_err_covdiag_mixed_vs_fp64": err_mixcov,
            "plain_vs_batched_max_rel_err": err_plain,
            "fp64_vs_oracle_max_rel_err": oracle_err,
            "kernels_eager_mixed": kern,
            "chunks": per_chunk,
            "mixed_chunks": mixed_chunks,
            "note": ("'plain': one ukhip_batch_create per problem (inputs resident), one smoothing per call, problems "
                     "one after another; 'batched': every problem of a chunk in one launch per level per state "
                     "dimension, dimensions on forked streams, replayed as a HIP graph; all estimates only except "
                     "'with_covariances'. Mixed covariances come from the FP32 factors (FP32 accuracy). 'mixed' "
                     "(the value) runs in chunks sized for the mixed path's own device memory (mixed_chunks; the "
                     "FP64 elimination's storage is never allocated); every other leg, and the errors, use the "
                     "FP64 chunking (chunks), where 'mixed_at_fp64_chunking' is the same path on those chunks."),
        },
    }


# ---------------------------------------------------------------------------------------------------------------
# secondary legs
# ---------------------------------------------------------------------------------------------------------------

def cpu_baseline(name, sample_k, algorithm="oddeven", want_outputs=False):
    """The reference itself (oracle/_ref: its C smoothers, sequential backend,
    one core) on a bounded sample of the workload; the oracle port if the
    reference build is absent. want_outputs: also return its estimates and
    covariance diagonals."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import uk_libs
    from ukhip import capi
    pr = make_problem(name, sample_k)
    ref = uk_libs.reference()
    kind = "reference"
    m0 = 0 if pr.o[0] is None else pr.o[0].shape[0]
    if algorithm == "associative" and m0 < pr.n[0]:
        ref = None  # the reference's associative prior needs m_0 >= n (it asserts otherwise)
    lib = ref
    if ref is None:
        lib = uk_libs.oracle()
        lib.lib.oracle_set_threads(1)
        kind = "port"
    eqs = capi.EquationSet(pr)
    fn = lib.oe if algorithm == "oddeven" else lib.asc
    t0 = time.perf_counter()
    fn(0, eqs.ptrs, pr.k)
    dt = time.perf_counter() - t0
    xs, cs = eqs.collect(lib.free, want_cov=want_outputs, read=want_outputs)
    why = " (a power of two: the reference odd-even aborts on other lengths)" if algorithm == "oddeven" else ""
    if kind == "port" and algorithm == "associative" and m0 < pr.n[0]:
        why += f"; the oracle's restatement, since the reference's associative smoother asserts when m_0={m0} < n"
    out = {"value": pr.k / dt, "unit": "timesteps/s", "cores": 1, "cpu": cpu_model(), "kind": kind,
           "sample": f"{WORKLOADS[name]['desc'].split(' k=')[0]} k={sample_k}{why}, "
                     f"kalman_smooth_{algorithm} with covariances, {dt:.1f} s"}
    if want_outputs:
        return out, pr, np.stack(xs), np.stack([np.diag(c) for c in cs])
    return out


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
    eqs.collect(o.free, read=False)
    o.lib.oracle_set_threads(1)
    return {"value": pr.k / dt, "unit": "timesteps/s", "cores": cores, "cpu": cpu_model(), "kind": "port",
            "sample": f"k={sample_k}, {dt:.1f} s"}


def mixed_legs(b, timesteps, x64, d64, args, label):
    """FP32 elimination + FP64 refinement against FP64 estimates-only on the
    same resident batch; errors of the estimates (x64) and of the covariance
    diagonals from the mixed factors (d64) against the FP64 results."""
    ms64 = _time(lambda: b.smooth("oddeven", options=NOCOV, sync=False), args.steps, 1)
    out = {"fp64_states_only": {"ms_per_step": ms64, "value": timesteps / (ms64 * 1e-3)}}
    for r in args.refinements:
        b.smooth_mixed_opt(r, 0, True)
        ex = max(step_rel_err(b.download(i, want_cov=False)[0], x64[i]) for i in range(len(x64)))
        ed = max(step_rel_err(b.download_cov(i, "D"), d64[i]) for i in range(len(d64)))
        ms = _time(lambda: b.smooth_mixed_opt(r, 0, False, sync=False), args.steps, 1)
        msc = _time(lambda: b.smooth_mixed_opt(r, 0, True, sync=False), 1, 0)
        out[f"refine{r}"] = {"ms_per_step": ms, "value": timesteps / (ms * 1e-3),
                             "speedup_vs_fp64_states_only": ms64 / ms, "max_rel_err_states": ex,
                             "ms_per_step_with_covariances": msc, "max_rel_err_covdiag": ed}
        say(f"{label} mixed, {r} refinement(s): {ms:.2f} ms/step ({ms64 / ms:.2f}x FP64 states-only), "
            f"err states {ex:.2e}, covariance diagonals {ed:.2e}")
    return out


def end_to_end(pr):
    """kalman_smooth_oddeven through the C ABI on the whole problem: host
    packing, upload, smoothing, download and the per-step output matrices
    (the step structures are built before the clock starts)."""
    from ukhip import capi
    L = device.lib()
    eqs = capi.EquationSet(pr)
    t0 = time.perf_counter()
    L.kalman_smooth_oddeven(0, eqs.ptrs, pr.k)
    dt = time.perf_counter() - t0
    err = device.last_error()
    eqs.collect(L.matrix_free, read=False)
    if err:
        raise RuntimeError(err)
    return {"ms": dt * 1e3, "value": pr.k / dt, "unit": "timesteps/s",
            "note": "kalman_smooth_oddeven(0, equations, k): packing, H2D, smoothing with covariances, D2H and "
                    "k state + k covariance matrices allocated with matrix_create"}


def run_long(args):
    """The single long problem (n=48, m=24, k=200000) in FP64 with per-kernel
    events (roofline of its dominant kernel), mixed precision, the associative
    smoother, the end-to-end C-ABI call, and errors against the reference."""
    name = "n48m24_k200000"
    t = time.perf_counter()
    pr = make_problem(name, args.k)
    say(f"generated {pr.name} k={pr.k} in {time.perf_counter() - t:.1f}s")
    b = device.Batch([pr])
    for _ in range(args.warmup):
        b.smooth("oddeven")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    kern = {nm: 0.0 for nm in device.OE_KERNELS}
    launches = None
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ms, launches = b.smooth_timed(0)
        for nm in kern:
            kern[nm] += ms[nm] / args.steps
    device.lib().ukhip_synchronize()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / args.steps
    say(f"n=48: {ms_per_step:.2f} ms/step; events: " + ", ".join(f"{k} {v:.3f}" for k, v in kern.items() if v))
    n, ml = 48, m_list_of(pr)
    fl = flops.oddeven_flops(n, ml, ["w"] * pr.k, ["w"] * pr.k, cov=True)
    fl_of = {"k_weigh": fl["weigh"], "k_phase1": fl["phase1"], "k_phase2": fl["phase2"], "k_phase3": fl["phase3"],
             "k_backsub": fl["backsub"]}
    dom = max(fl_of, key=lambda k: kern[k])
    label = kernel_label(dom, n, 24)
    achieved = fl_of[dom] / (kern[dom] * 1e-3) / 1e12
    traffic = traffic_of(label.replace(", *>", ">").rsplit(", ", 2)[0] + ">" if dom.startswith("k_phase") else label,
                         name if args.k is None else None)
    x64, S64 = b.download(0)
    d64 = np.einsum("kii->ki", S64)
    out = {"workload": WORKLOADS[name]["desc"] + ("" if args.k is None else f" (k = {pr.k})"),
           "ms_per_step": ms_per_step, "value": pr.k / (ms_per_step * 1e-3),
           "roofline": {"bound": "valu" if "mfma" not in label else "mfma", "kernel": label, "achieved": achieved,
                        "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                        "traffic": traffic, "algorithmic_flops_per_step": fl_of[dom],
                        "launches_per_step": launches[dom], "kernel_ms_per_step": kern[dom],
                        "avg_launch_ms": kern[dom] / launches[dom]},
           "kernels_ms_per_step": {k: v for k, v in kern.items() if v},
           "kernels_tflops": {k: fl_of[k] / (kern[k] * 1e-3) / 1e12 for k in fl_of if kern[k] > 0},
           "algorithmic_flops_per_timestep": fl["total"] / pr.k}
    del S64
    # the diagonal-only covariance path on the same resident batch
    dms = _time(lambda: b.smooth("oddeven", options=COVDIAG, sync=False), args.steps, 1)
    xd, dd = b.download(0, want_cov=False)[0], b.download_cov(0, "D")
    out["diagonal_only"] = {"ms_per_step": dms, "value": pr.k / (dms * 1e-3), "unit": "timesteps/s",
                            "speedup_vs_full_covariances": ms_per_step / dms,
                            "states_bit_identical_to_full": bool(np.array_equal(xd, x64)),
                            "max_rel_err_covdiag_vs_full": step_rel_err(dd, d64)}
    say(f"n=48 diagonal-only covariances: {dms:.2f} ms/step, diagonals vs full {step_rel_err(dd, d64):.1e}")
    if not args.no_mixed:
        out["mixed_precision"] = mixed_legs(b, pr.k, [x64], [d64], args, "n=48")
    if not args.no_associative:
        for _ in range(max(1, args.warmup // 2)):
            b.smooth("associative")
        ta = time.perf_counter()
        for _ in range(args.assoc_steps):
            b.smooth("associative", sync=False)
        device.lib().ukhip_synchronize()
        ams = (time.perf_counter() - ta) * 1e3 / args.assoc_steps
        xa, Sa = b.download(0)
        # per kernel class (HIP events after each class), and the roofline of the dominant one
        ams_k, anl = b.smooth_associative_timed()
        afl = flops.associative_flops(pr.n[0], m_list_of(pr))
        cls = [c for c in device.AS_KERNELS if c != "total"]
        adom = max(cls, key=lambda c: ams_k[c])
        a_ach = afl[adom] / (ams_k[adom] * 1e-3) / 1e12
        names = {"fscan": "ka_scan_reduce/round/down<48, true>", "sscan": "ka_scan_reduce/round/down<48, false>",
                 "felem": "ka_felem<48>", "selem": "ka_selem<48>", "model": "ka_model<48>", "prior": "ka_prior<48>",
                 "filtered": "ka_filtered", "results": "ka_results", "weigh": "k_weigh<48>"}
        out["associative"] = {
            "algorithm": "associative scan (Saerkkae-Garcia-Fernandez), FP64, with covariances",
            "ms_per_step": ams, "value": pr.k / (ams * 1e-3), "unit": "timesteps/s", "steps": args.assoc_steps,
            "max_rel_err_states_vs_oddeven": step_rel_err(xa, x64),
            "max_rel_err_covdiag_vs_oddeven": step_rel_err(np.einsum("kii->ki", Sa), d64),
            "kernels_ms_per_step": {c: ams_k[c] for c in device.AS_KERNELS},
            "kernels_tflops": {c: afl[c] / (ams_k[c] * 1e-3) / 1e12 for c in cls if afl[c] > 0 and ams_k[c] > 0},
            "roofline": {"bound": "valu", "kernel": names[adom], "achieved": a_ach, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": a_ach / FP64_PEAK_TFLOPS, "traffic": None,
                         "algorithmic_flops_per_step": afl[adom], "launches_per_step": anl[adom],
                         "kernel_ms_per_step": ams_k[adom], "avg_launch_ms": ams_k[adom] / max(anl[adom], 1),
                         "measured": "HIP events after each kernel class of one associative smoothing",
                         "flops": "ukhip.flops.associative_flops: the reference's matrix operations, two "
                                  "combines per element in each scan"},
            "algorithmic_tflops_whole_path": afl["total"] / (ams * 1e-3) / 1e12}
        del Sa
        say(f"associative {ams:.2f} ms/step, vs odd-even: states {out['associative']['max_rel_err_states_vs_oddeven']:.2e}"
            f", covariance diagonals {out['associative']['max_rel_err_covdiag_vs_oddeven']:.2e}")
    b.free()
    if not args.no_e2e:
        out["end_to_end_c_abi"] = end_to_end(pr)
        say(f"end to end through kalman_smooth_oddeven: {out['end_to_end_c_abi']['ms']:.0f} ms")
    del pr
    if not args.no_cpu_baseline:
        try:
            base, prs, rx, rd = cpu_baseline(name, args.cpu_sample_k, want_outputs=True)
            out["cpu_baseline"] = base
            bs = device.Batch([prs])
            bs.smooth("oddeven")
            gx, gS = bs.download(0)
            bs.free()
            out["vs_reference_output"] = {
                "sample": f"the same workload at k={args.cpu_sample_k}, GPU odd-even against the reference's own "
                          "kalman_smooth_oddeven output (oracle/_ref)" if base["kind"] == "reference" else
                          "against the oracle (no reference build)",
                "max_rel_err_states": step_rel_err(gx, rx),
                "max_rel_err_covdiag": step_rel_err(np.einsum("kii->ki", gS), rd)}
            out["cpu_port_allcores"] = cpu_port_allcores(name, args.cpu_sample_k)
            if not args.no_associative:
                abase, aprs, arx, ard = cpu_baseline(name, args.cpu_sample_k_assoc, "associative", want_outputs=True)
                out["associative"]["cpu_baseline"] = abase
                ba = device.Batch([aprs])
                ba.smooth("associative")
                ax, aS = ba.download(0)
                ba.free()
                out["associative"]["vs_cpu_output"] = {
                    "sample": f"the same workload at k={args.cpu_sample_k_assoc}, GPU associative against the "
                              f"{'reference' if abase['kind'] == 'reference' else 'oracle'}'s associative output "
                              "on the same inputs",
                    "max_rel_err_states": step_rel_err(ax, arx),
                    "max_rel_err_covdiag": step_rel_err(np.einsum("kii->ki", aS), ard)}
        except Exception as e:  # the GPU figures stand without it
            out["cpu_baseline"] = {"error": str(e)}
    return out


def precision_workload(args):
    """The ill-conditioned precision workload (n=64, m=32, k=200000, F = Q
    diag(logspace(0,-3,64)), K = 1e-2 I, C = I): FP64 with covariances, FP64
    estimates only, and FP32 + FP64 refinement, errors on the estimates and on
    the covariance diagonals against the FP64 results."""
    pr = make_problem("n64m32_k200000", args.k, lists=False)
    b = device.Batch([pr])
    ms_cov = _time(lambda: b.smooth("oddeven", sync=False), args.steps, 1)
    x64, S64 = b.download(0)
    d64 = np.einsum("kii->ki", S64)
    del S64
    res = {"workload": WORKLOADS["n64m32_k200000"]["desc"] + ("" if args.k is None else f" (k = {pr.k})"),
           "fp64_with_covariances": {"ms_per_step": ms_cov, "value": pr.k / (ms_cov * 1e-3)}}
    res.update(mixed_legs(b, pr.k, [x64], [d64], args, "n=64"))
    b.free()
    return res


def run_serving(args):
    """A bounded sample of the serving workload: --serve-requests requests (n
    uniform over {3..16}, m uniform over [1,n], lengths log-uniform in
    [100, 50000], F = 0.99 Q_n, G ~ N(0,1/n), K = I, C = 0.1 I, 5% of steps
    unobserved), drawn in turn from a pool of --serve-pool generated ones. For
    each mode (FP64 with covariances; FP32 + refinement, estimates) the stream
    is submitted at once (capacity), then as a Poisson stream at --serve-load
    times the FP64 capacity (the same arrival times for both modes). Latency is
    the server's own, from submit to results ready on the host."""
    import threading
    specs = P.serve_specs(args.serve_requests, seed=5)
    t = time.perf_counter()
    pool = [P.serve_problem(s, lists=False) for s in specs[:args.serve_pool]]
    reqs = [pool[i % len(pool)] for i in range(args.serve_requests)]
    steps = sum(p.k for p in reqs)
    say(f"serving: {len(reqs)} requests ({steps} steps) from a pool of {len(pool)}, generated in "
        f"{time.perf_counter() - t:.1f}s")

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

    out = {"sample": f"{len(reqs)} requests drawn from {len(pool)} generated ones (the configuration's 50,000 "
                     f"requests hold ~4e8 steps, ~1 TB of inputs: beyond a bounded run)",
           "requests": len(reqs), "pool": len(pool), "timesteps": steps, "hbm_budget_gb": args.serve_budget_gb,
           "max_batch": args.serve_max_batch, "max_inflight": args.serve_inflight}
    modes = (("fp64", -1), (f"mixed_refine{args.serve_refine}", args.serve_refine))
    for name, refine in modes:
        out[name] = {"all_at_once": stream(refine, None)}
        say(f"serving {name}, all at once: {out[name]['all_at_once']}")
    lam = args.serve_load * out["fp64"]["all_at_once"]["requests_per_s"]
    gaps = P.rng_for(77).exponential(1.0 / lam, len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    out["poisson_rate_requests_per_s"] = lam
    out["offered_timesteps_per_s"] = lam * steps / len(reqs)
    for name, refine in modes:
        out[name]["poisson"] = stream(refine, arrivals)
        say(f"serving {name}, Poisson: {out[name]['poisson']}")
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
    say(f"filter: {nf} filters n={n} m={m}, {args.filter_steps} steps in {wall:.3f} s")
    # the same steps with the inputs already in device memory (ukhip_filter_step_device)
    ups = [device.Filter.upload(F=F, K=K, C=Cw, **steps[i]) for i in range(8)]
    flt = device.Filter(nf, n, m)
    for i in range(2):
        flt.step_device(ups[i % 8])
    device.lib().ukhip_synchronize()
    t0 = time.perf_counter()
    for i in range(args.filter_steps):
        flt.step_device(ups[i % 8])
    device.lib().ukhip_synchronize()
    wall_d = time.perf_counter() - t0
    xd, _ = flt.estimates(want_R=False)
    flt.close()
    for u in ups:
        for a in u.values():
            if isinstance(a, device.DeviceArray):
                a.free()
    assert np.all(np.isfinite(xd))
    say(f"filter, device-resident inputs: {args.filter_steps} steps in {wall_d:.3f} s")
    return {"filters": nf, "n": n, "m": m, "steps": args.filter_steps, "ms_per_step": wall * 1e3 / args.filter_steps,
            "value": nf * args.filter_steps / wall, "unit": "filter-steps/s",
            "note": "host inputs uploaded every step (PCIe included)",
            "device_inputs": {"ms_per_step": wall_d * 1e3 / args.filter_steps,
                              "value": nf * args.filter_steps / wall_d, "unit": "filter-steps/s",
                              "note": "inputs already in device memory (ukhip_filter_step_device), eight step "
                                      "input sets reused in turn"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch-count", type=int, default=4096, help="problems of the batching workload")
    ap.add_argument("--hbm-budget-gb", type=float, default=100.0,
                    help="device memory per resident chunk of the batching workload (FP64 legs)")
    ap.add_argument("--mixed-hbm-budget-gb", type=float, default=200.0,
                    help="device memory per resident chunk of the batching workload's mixed-precision leg")
    ap.add_argument("--refine", type=int, default=1, help="headline: FP64 refinement steps after the FP32 elimination")
    ap.add_argument("--k", type=int, default=None, help="override the long workloads' number of steps (testing only)")
    ap.add_argument("--no-headline", action="store_true")
    ap.add_argument("--no-long", action="store_true", help="skip the n=48 k=200000 workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-sample", action="store_true")
    ap.add_argument("--cpu-sample-k", type=int, default=8192)
    ap.add_argument("--no-associative", action="store_true")
    ap.add_argument("--assoc-steps", type=int, default=2)
    ap.add_argument("--cpu-sample-k-assoc", type=int, default=2048)
    ap.add_argument("--refinements", type=lambda s: [int(v) for v in s.split(",")], default=[1, 2],
                    help="mixed-precision legs of the long workloads: FP64 refinement steps")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the n=64 precision workload")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-filter", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in QR launch table")
    ap.add_argument("--launch-table", default=None, metavar="JSON",
                    help='set QR launch-table entries instead of autotuning, e.g. \'{"phase2<48, double>": "c2w2"}\'')
    ap.add_argument("--filter-count", type=int, default=4096)
    ap.add_argument("--filter-n", type=int, default=16)
    ap.add_argument("--filter-steps", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--serve-pool", type=int, default=300)
    ap.add_argument("--serve-budget-gb", type=float, default=32.0)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-refine", type=int, default=2)
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="Poisson rate as a fraction of the FP64 server's all-at-once request rate")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the long workload's timed region, copy MIB MiB device-to-device once (a plain copy "
                         "kernel of known size for calibrating rocprofv3's FETCH_SIZE and WRITE_SIZE counters)")
    ap.add_argument("--headline-mixed-only", action="store_true",
                    help="profiling aid: the headline runs its timed mixed-precision leg only (no FP64 legs, "
                         "ratios or errors)")
    ap.add_argument("--mixed-parts-only", action="store_true",
                    help="profiling aid: with --headline-mixed-only, the mixed leg runs only the eager smoothing per "
                         "chunk (parts one after another) whose per-launch times give the roofline; ms_per_step is "
                         "then that smoothing's summed device time")
    ap.add_argument("--no-graphs", action="store_true",
                    help="profiling aid: launch kernel by kernel everywhere (no HIP graph capture or replay)")
    ap.add_argument("--mixed-chunk-only", type=int, default=0, metavar="N",
                    help="profiling aid: the first N problems of the batching workload as one device batch, two eager "
                         "mixed smoothings (graphs off; --copy-calibration runs between them), no bench line")
    args = ap.parse_args()
    if args.mixed_parts_only and not args.headline_mixed_only:
        ap.error("--mixed-parts-only times the mixed leg alone: give --headline-mixed-only too")
    if args.no_graphs:
        device.lib().ukhip_use_graphs(0)
    if args.mixed_chunk_only:
        probs = [P.batch_problem(sp, lists=False) for sp in P.batch_specs(count=args.mixed_chunk_only)]
        b = device.Batch(probs)
        L = device.lib()
        L.ukhip_use_graphs(0)
        b.smooth_mixed_opt(args.refine, 0, False)
        if args.copy_calibration:
            copy_calibration(args.copy_calibration)
        b.smooth_mixed_opt(args.refine, 0, False)
        L.ukhip_synchronize()
        b.free()
        say(f"mixed chunk of {len(probs)} problems: two eager smoothings done")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box here has one MI355X); --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    table_src = "built-in"
    if args.launch_table:
        device.set_launch_table(json.loads(args.launch_table))
        table_src = "set by --launch-table"
    elif not args.no_autotune:
        t = time.perf_counter()
        device.autotune(64, 65536)
        table_src = "autotuned (ukhip_autotune)"
        say(f"autotuned the QR launch table in {time.perf_counter() - t:.1f}s: {device.launch_table()}")

    out = {"metric": METRIC, "value": None, "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 elimination + f64 refinement",
           "data": "synthetic (seeded Philox inputs, i.i.d. N(0,1) observations)",
           "config": {"workload": "batching: " + BATCH_DESC,
                      "algorithm": f"odd-even, FP32 elimination + {args.refine} FP64 refinement step(s) of the "
                                   "estimates, batched (one launch per level per state dimension, dimensions on "
                                   + ("one after another, eager (--mixed-parts-only))" if args.mixed_parts_only
                                      else "forked streams, kernel by kernel (--no-graphs))" if args.no_graphs else
                                      "forked streams, one HIP graph replay per chunk)")}}
    if not args.no_headline:
        h = run_headline(args)
        out["value"] = h["value"]
        out["ms_per_step"] = h["ms_per_step"]
        out["config"]["timesteps"] = h["total_steps"]
        out["config"]["hbm_budget_gb"] = args.hbm_budget_gb
        out["roofline"] = h["roofline"]
        out["headline"] = h["headline"]
    if not args.no_long:
        out["long_n48"] = run_long(args)
        if "cpu_baseline" in out["long_n48"]:
            out["cpu_baseline"] = out["long_n48"]["cpu_baseline"]
            out["cpu_port_allcores"] = out["long_n48"].get("cpu_port_allcores")
    if not args.no_mixed and not args.no_precision:
        out["precision_workload"] = precision_workload(args)
    if not args.no_serving:
        out["serving"] = run_serving(args)
    if not args.no_filter:
        out["sequential_filter"] = run_filter(args)
    out["qr_launch_table"] = device.launch_table()
    out["qr_launch_table_source"] = table_src
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
