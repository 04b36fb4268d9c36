"""Writes the golden fixtures in this directory from the reference's own C smoothers.

Run in a container holding the reference sources (/root/reference) after
`make -C oracle ref` has built oracle/_ref/libukref.so from them:

    python tests/golden/make_golden.py                 # every fixture
    python tests/golden/make_golden.py --only-shapes   # shapes_oddeven.npz only

Every reference call runs in a forked child, so a reference assertion
(its defects D1-D5) costs one case, not the script. Each fixture holds inputs
(explicit for small cases; generator spec + SHA-256 for large ones, whose
inputs are regenerated bit-identically from Philox streams) and the
reference's outputs.
"""
import json
import multiprocessing as mp
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import uk_libs  # noqa: E402
from ukhip import capi  # noqa: E402
from ukhip import problems as P  # noqa: E402

PERF_C = "/root/reference/c/performance.c"


def _child(conn, algo, problem, options):
    try:
        ref = uk_libs.reference()
        fn = ref.oe if algo == "oddeven" else ref.asc
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)  # the reference prints on create; keep stdout clean
        xs, cs = capi.run(fn, ref.free, problem, options)
        conn.send(("ok", xs, cs))
    except Exception as e:  # pragma: no cover
        conn.send(("err", str(e), None))
    conn.close()


def run_ref(algo, problem, options=0):
    ctx = mp.get_context("fork")
    a, b = ctx.Pipe()
    p = ctx.Process(target=_child, args=(b, algo, problem, options))
    p.start()
    b.close()  # the parent's copy: a dead child then reads as EOF at once
    try:
        msg = a.recv() if a.poll(3600) else ("err", "timeout", None)
    except EOFError:
        msg = ("err", "reference aborted", None)
    p.join(60)
    if p.is_alive():
        p.kill()
        p.join()
    return None if msg[0] != "ok" else (msg[1], msg[2])


def pack_outputs(xs, cs, every=1, full_cov=True):
    idx = np.arange(0, len(xs), every)
    if len(xs) - 1 not in idx:
        idx = np.append(idx, len(xs) - 1)
    out = dict(out_index=idx.astype(np.int64),
               out_state=np.stack([xs[i] for i in idx]),
               out_state_maxabs=np.max(np.abs(np.stack(xs)), axis=0),
               out_state_sumsq=np.sum(np.stack(xs) ** 2, axis=0))
    if cs[0] is not None:
        full = [c for c in cs]
        out["out_covdiag"] = np.stack([np.diag(full[i]) if full[i] is not None else
                                       np.full(len(xs[0]), np.nan) for i in idx])
        if every == 1 and full_cov:
            out["out_cov"] = np.stack([c if c is not None else np.full((len(xs[0]),) * 2, np.nan)
                                       for c in full])
    return out


def parse_perf_arrays():
    src = open(PERF_C).read()
    arrs = {}
    for name in ("F6", "G6", "o6", "F48", "G48", "o48"):
        m = re.search(r"double\s+%s\[\]\s*=\s*\{([^}]*)\}" % name, src)
        arrs[name] = np.array([float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()])
    return arrs


def perf_problem(arrs, n, k):
    """performance.c:328-340: H=I, F=F<n>, c=0, K=I('W'), G=G<n>, o=o<n>, C=I('W')
    for every step (matrix_create_from_rowwise => row-major data)."""
    F = arrs["F%d" % n].reshape(n, n)
    G = arrs["G%d" % n].reshape(n, n)
    o = arrs["o%d" % n].reshape(n, 1)
    I = np.eye(n)
    return P.Problem(k=k, n=[n] * k, H=[None] + [I] * (k - 1), F=[None] + [F] * (k - 1),
                     c=[None] + [np.zeros((n, 1))] * (k - 1), K=[None] + [I] * (k - 1),
                     K_type=[None] + ["W"] * (k - 1), G=[G] * k, o=[o] * k, C=[I] * k, C_type=["W"] * k,
                     name=f"performance_n{n}_k{k}")


def save(name, arrays, meta):
    arrays = dict(arrays)
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print("wrote", name, meta.get("note", ""), flush=True)


def small_cases():
    kinds = ["W", "w", "C", "U", "F"]
    cases = []
    seed = 1000
    for n in (1, 2, 3, 6):
        for k in (1, 2, 4, 8, 16, 64):
            for variant in ("m=n", "m<n", "gaps", "m>n"):
                seed += 1
                m = {"m=n": n, "m<n": max(1, n - 1), "gaps": n, "m>n": n + 2}[variant]
                unob = ()
                if variant == "gaps" and k >= 4:
                    unob = tuple(range(1, k - 1, 3))
                cases.append(dict(seed=seed, k=k, n=n, m=m, unobserved=unob,
                                  K_kind=kinds[seed % 5], C_kind=kinds[(seed // 5) % 5],
                                  H_kind="R" if seed % 3 == 0 else "I"))
    return cases


def shapes_fixture():
    """Per-step shapes of the reference's test catalogue
    (matlab/examples/kalmanTest.m:80-95: rectangular H, changing state
    dimension, more observation rows than states) through the reference's
    kalman_smooth_oddeven. Cases whose length is not a power of two abort in
    the reference (defect D1) and are recorded as such, inputs kept."""
    kinds = [("W", "W"), ("C", "w"), ("w", "C"), ("U", "W")]
    bundle, ok, aborted = {}, [], []
    for t, (name, rows) in enumerate(sorted(P.shape_catalogue().items())):
        K_kind, C_kind = kinds[t % len(kinds)]
        pr = P.shape_problem(5000 + t, rows, K_kind=K_kind, C_kind=C_kind)
        tag = f"{name}_{K_kind}{C_kind}"
        arrs = P.problem_to_arrays(pr)
        r = run_ref("oddeven", pr)
        for key, v in arrs.items():
            bundle[tag + "__" + key] = v
        if r is None:
            aborted.append(tag)
            continue
        ok.append(tag)
        xs, cs = r
        # per-step dimensions differ: states and covariances flattened step after step
        bundle[tag + "__out_state_flat"] = np.concatenate([np.ravel(x) for x in xs])
        bundle[tag + "__out_cov_flat"] = np.concatenate([np.ravel(c, order="F") for c in cs])
    save("shapes_oddeven", bundle, dict(cases=ok + aborted, reference_ok=ok, reference_aborted=aborted,
                                        note="reference kalman_smooth_oddeven on kalmanTest.m shape tables"))


def main():
    if uk_libs.reference() is None:
        sys.exit("oracle/_ref/libukref.so is missing: run `make -C oracle ref` first")
    os.makedirs(HERE, exist_ok=True)
    if "--only-shapes" in sys.argv:
        shapes_fixture()
        return
    shapes_fixture()

    # 1. small odd-even and associative cases, inputs stored explicitly
    oe_ok, oe_fail, as_ok, as_fail = [], [], [], []
    bundle_oe, bundle_as = {}, {}
    for c in small_cases():
        spec = dict(c)
        pr = P.random_problem(spec.pop("seed"), spec.pop("k"), spec.pop("n"), spec.pop("m"), **spec)
        tag = f"s{c['seed']}_n{c['n']}_k{c['k']}_m{c['m']}_{c['K_kind']}{c['C_kind']}{c['H_kind']}_u{len(c['unobserved'])}"
        arrs = P.problem_to_arrays(pr)
        r = run_ref("oddeven", pr)
        if r is None:
            oe_fail.append(tag)
        else:
            oe_ok.append(tag)
            for key, v in {**arrs, **pack_outputs(*r)}.items():
                bundle_oe[tag + "__" + key] = v
        all_obs = all(o is not None for o in pr.o)
        if c["H_kind"] == "I" and all_obs and c["m"] >= c["n"] and c["k"] >= 2:
            r = run_ref("associative", pr)
            if r is None:
                as_fail.append(tag)
            else:
                as_ok.append(tag)
                for key, v in {**arrs, **pack_outputs(*r)}.items():
                    bundle_as[tag + "__" + key] = v
    save("small_oddeven", bundle_oe, dict(cases=oe_ok, reference_aborted=oe_fail,
                                           note="reference kalman_smooth_oddeven, small seeded problems"))
    save("small_associative", bundle_as, dict(cases=as_ok, reference_aborted=as_fail,
                                               note="reference kalman_smooth_associative, all observed, m0>=n, H=I"))

    # 2. the fixed inputs of the reference's own performance driver
    arrs = parse_perf_arrays()
    for n, k in ((6, 1024), (48, 128)):
        pr = perf_problem(arrs, n, k)
        base = dict(perf_F=arrs["F%d" % n], perf_G=arrs["G%d" % n], perf_o=arrs["o%d" % n])
        for algo in ("oddeven", "associative"):
            r = run_ref(algo, pr)
            assert r is not None, (algo, n, k)
            outs = pack_outputs(*r, every=1, full_cov=(n == 6))
            sel = np.array([0, 1, k // 2, k - 1])
            outs["out_cov_sel_index"] = sel
            outs["out_cov_sel"] = np.stack([r[1][i] if r[1][i] is not None else np.full((n, n), np.nan)
                                            for i in sel])  # D4: NULL last covariance
            save(f"performance_n{n}_k{k}_{algo}", {**base, **outs},
                 dict(n=n, k=k, algo=algo, note="c/performance.c inputs (F%d,G%d,o%d)" % (n, n, n)))

    # 3. large configurations, regenerated from Philox seeds (SHA-256 checked)
    F6 = P.random_orthogonal(P.rng_for(1001), 6)
    F48 = 0.99 * P.random_orthogonal(P.rng_for(1002), 48)
    F64 = P.random_orthogonal(P.rng_for(1004), 64) @ np.diag(np.logspace(0, -3, 64))
    large = [
        ("rot6_k131072_oddeven", "oddeven", dict(k=131072, n=6, m=6, seed=11, F=F6, K_kind="W", C_kind="W"), 1024),
        ("rot6_k100000_associative", "associative", dict(k=100000, n=6, m=6, seed=12, F=F6, K_kind="W", C_kind="W"), 1000),
        ("n48m24_k4096_oddeven", "oddeven", dict(k=4096, n=48, m=24, seed=13, F=F48, G_scale=1.0 / np.sqrt(48.0),
                                                 Kvar=1.0, Cvar=0.1, unob10=True), 64),
        ("n64m32_k2048_oddeven", "oddeven", dict(k=2048, n=64, m=32, seed=14, F=F64, G_scale=1.0 / 8.0,
                                                 Kvar=1e-2, Cvar=1.0), 32),
    ]
    for name, algo, spec, every in large:
        spec = dict(spec)
        F = spec["F"]
        unob10 = spec.pop("unob10", False)
        pr = P.iid_problem(spec.pop("k"), spec.pop("n"), spec.pop("m"),
                           unobserved=(lambda i: i % 10 == 9) if unob10 else None, **spec)
        digest = P.input_digest(pr)
        r = run_ref(algo, pr)
        assert r is not None, name
        gen = {kk: (vv if not isinstance(vv, np.ndarray) else None) for kk, vv in spec.items()}
        gen.pop("F")
        save(name, {"gen_F": F, **pack_outputs(*r, every=every)},
             dict(algo=algo, k=pr.k, n=pr.n[0], m=int(pr.packed["m"]), unob10=unob10, gen=gen,
                  sha256=digest, note="iid_problem regenerated from Philox; F stored"))


if __name__ == "__main__":
    main()
