"""Pins the CPU oracle (oracle/uk_oracle.c) before it is trusted as the checker.

- golden fixtures written by the reference's own C smoothers (tests/golden);
- the global least-squares definition (matlab/KalmanSparse.m:142-306);
- closed forms: the n=6 rotation problem's smoothed covariances and the
  scalar textbook Kalman/RTS smoother;
- the LAWN-41 FLOP replay against the reference's measured call shapes.
"""
import json
import os

import numpy as np
import pytest

import uk_libs
from checks import max_rel_err, rel_err
from ukhip import problems as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# Agreement bars against the reference: both are FP64 QR-based; observed
# differences are ~1e-14. The associative algorithm works in covariance form
# and the reference itself departs from the global LS by up to ~1e-8 on some
# random problems, hence its looser bar.
TOL_OE = 1e-9
TOL_ASSOC = 1e-7


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(bytes(d["meta"]).decode())
    return d, meta


def sub_arrays(d, tag):
    pre = tag + "__"
    return {k[len(pre):]: d[k] for k in d.files if k.startswith(pre)}


@pytest.mark.parametrize("fixture,algo,tol", [("small_oddeven", "oddeven", TOL_OE),
                                               ("small_associative", "associative", TOL_ASSOC)])
def test_small_fixtures(oracle, fixture, algo, tol):
    d, meta = load(fixture)
    assert len(meta["cases"]) >= 20
    for tag in meta["cases"]:
        a = sub_arrays(d, tag)
        pr = P.problem_from_arrays(a, tag)
        xs, cs = getattr(oracle, algo)(pr)
        ref_x = list(a["out_state"])
        ref_c = list(a["out_cov"])
        skip = (pr.k - 1,) if algo == "associative" else ()  # D4: reference last covariance is NULL
        assert max_rel_err(xs, ref_x) < tol, tag
        assert max_rel_err(cs, ref_c, skip=skip) < tol, tag


@pytest.mark.parametrize("n,k,algo", [(6, 1024, "oddeven"), (6, 1024, "associative"),
                                      (48, 128, "oddeven"), (48, 128, "associative")])
def test_performance_driver_inputs(oracle, n, k, algo):
    d, meta = load(f"performance_n{n}_k{k}_{algo}")
    F = d["perf_F"].reshape(n, n)
    G = d["perf_G"].reshape(n, n)
    o = d["perf_o"].reshape(n, 1)
    I = np.eye(n)
    pr = P.Problem(k=k, n=[n] * k, H=[None] + [I] * (k - 1), F=[None] + [F] * (k - 1),
                   c=[None] + [np.zeros((n, 1))] * (k - 1), K=[None] + [I] * (k - 1),
                   K_type=[None] + ["W"] * (k - 1), G=[G] * k, o=[o] * k, C=[I] * k, C_type=["W"] * k)
    xs, cs = getattr(oracle, algo)(pr)
    tol = TOL_OE if algo == "oddeven" else TOL_ASSOC
    assert max_rel_err(xs, list(d["out_state"])) < tol
    diag = np.stack([np.diag(c) for c in cs])
    last = k - 1 if algo == "associative" else k
    assert rel_err(diag[:last], d["out_covdiag"][:last]) < tol
    for j, i in enumerate(d["out_cov_sel_index"]):
        if algo == "associative" and i == k - 1:
            continue
        assert rel_err(cs[i], d["out_cov_sel"][j]) < tol


def _large_problem(meta, F):
    g = dict(meta["gen"])
    unob = (lambda i: i % 10 == 9) if meta["unob10"] else None
    return P.iid_problem(meta["k"], meta["n"], meta["m"], F=F, unobserved=unob, **g)


@pytest.mark.slow
@pytest.mark.parametrize("name", ["rot6_k131072_oddeven", "rot6_k100000_associative",
                                  "n48m24_k4096_oddeven", "n64m32_k2048_oddeven"])
def test_large_fixtures(oracle, name):
    d, meta = load(name)
    pr = _large_problem(meta, d["gen_F"])
    if P.input_digest(pr) != meta["sha256"]:
        pytest.skip("regenerated inputs differ from the fixture's (numpy Philox stream changed)")
    oracle.lib.oracle_set_threads(os.cpu_count() or 1)
    try:
        xs, cs = getattr(oracle, meta["algo"])(pr)
    finally:
        oracle.lib.oracle_set_threads(1)
    idx = d["out_index"]
    tol = TOL_OE if meta["algo"] == "oddeven" else TOL_ASSOC
    assert max_rel_err([xs[i] for i in idx], list(d["out_state"])) < tol
    X = np.stack(xs)
    assert rel_err(np.max(np.abs(X), axis=0), d["out_state_maxabs"]) < tol
    assert rel_err(np.sum(X ** 2, axis=0), d["out_state_sumsq"]) < tol
    diag = np.stack([np.diag(cs[i]) for i in idx])
    ref = d["out_covdiag"]
    keep = idx != pr.k - 1 if meta["algo"] == "associative" else slice(None)
    assert rel_err(diag[keep], ref[keep]) < tol


def test_dense_ls_definition(oracle):
    cases = [dict(k=9, n=4, m=2), dict(k=10, n=4, m=4, unobserved=(0, 3, 4, 9)),
             dict(k=8, n=3, m=3, K_kind="C", C_kind="w", H_kind="R"), dict(k=7, n=5, m=1),
             dict(k=6, n=3, m=5, K_kind="U", C_kind="F"), dict(k=27, n=1, m=2, unobserved=(9, 21, 24, 25))]
    for s, c in enumerate(cases):
        c = dict(c)
        pr = P.random_problem(500 + s, c.pop("k"), c.pop("n"), c.pop("m"), **c)
        lx, lc = uk_libs.dense_ls(pr)
        xs, cs = oracle.oddeven(pr)
        assert max_rel_err(xs, lx) < 1e-12 and max_rel_err(cs, lc) < 1e-12, c
        ax, ac = oracle.associative(pr)
        assert max_rel_err(ax, lx) < 1e-9 and max_rel_err(ac, lc) < 1e-9, c


def test_rotation_closed_form(oracle):
    """F orthogonal, H=G=K=C=I: interior smoothed covariance (1/sqrt 5) I,
    first and last ((sqrt 5 - 1)/2) I (the stationary two-sided Riccati fixed point)."""
    k = 257
    F = P.random_orthogonal(P.rng_for(77), 6)
    pr = P.iid_problem(k, 6, 6, seed=78, F=F, K_kind="W", C_kind="W")
    for algo in ("oddeven", "associative"):
        xs, cs = getattr(oracle, algo)(pr)
        mid = cs[k // 2]
        assert np.max(np.abs(mid - np.eye(6) / np.sqrt(5))) < 1e-13, algo
        end = (np.sqrt(5) - 1) / 2
        assert np.max(np.abs(cs[0] - end * np.eye(6))) < 1e-13, algo
        assert np.max(np.abs(cs[-1] - end * np.eye(6))) < 1e-13, algo


def _scalar_rts(F, Q, G, R, o, c):
    """Textbook scalar Kalman filter + Rauch-Tung-Striebel smoother, flat prior."""
    k = len(o)
    xf = np.zeros(k)
    Pf = np.zeros(k)
    xf[0] = o[0] / G[0]
    Pf[0] = R[0] / G[0] ** 2
    for i in range(1, k):
        xp = F[i] * xf[i - 1] + c[i]
        Pp = F[i] ** 2 * Pf[i - 1] + Q[i]
        Kg = Pp * G[i] / (G[i] ** 2 * Pp + R[i])
        xf[i] = xp + Kg * (o[i] - G[i] * xp)
        Pf[i] = (1 - Kg * G[i]) * Pp
    xs, Ps = xf.copy(), Pf.copy()
    for i in range(k - 2, -1, -1):
        Pp = F[i + 1] ** 2 * Pf[i] + Q[i + 1]
        E = Pf[i] * F[i + 1] / Pp
        xs[i] = xf[i] + E * (xs[i + 1] - (F[i + 1] * xf[i] + c[i + 1]))
        Ps[i] = Pf[i] + E * (Ps[i + 1] - Pp) * E
    return xs, Ps


def test_scalar_textbook_rts(oracle):
    rng = P.rng_for(91)
    k = 37
    F = rng.uniform(-1.2, 1.2, k)
    Q = rng.uniform(0.1, 2.0, k)
    G = rng.uniform(0.5, 2.0, k)
    R = rng.uniform(0.1, 2.0, k)
    o = rng.standard_normal(k)
    c = rng.standard_normal(k)
    pr = P.Problem(k=k, n=[1] * k, H=[None] + [np.eye(1)] * (k - 1),
                   F=[None] + [np.array([[F[i]]]) for i in range(1, k)],
                   c=[None] + [np.array([[c[i]]]) for i in range(1, k)],
                   K=[None] + [np.array([[Q[i]]]) for i in range(1, k)], K_type=[None] + ["C"] * (k - 1),
                   G=[np.array([[G[i]]]) for i in range(k)], o=[np.array([[o[i]]]) for i in range(k)],
                   C=[np.array([[R[i]]]) for i in range(k)], C_type=["C"] * k)
    xt, Pt = _scalar_rts(F, Q, G, R, o, c)
    for algo in ("oddeven", "associative"):
        xs, cs = getattr(oracle, algo)(pr)
        assert np.max(np.abs(np.array([x[0] for x in xs]) - xt) / np.abs(xt).max()) < 1e-12, algo
        assert np.max(np.abs(np.array([v[0, 0] for v in cs]) - Pt) / Pt.max()) < 1e-12, algo


@pytest.mark.parametrize("n,k,nocov,per_step", [(6, 1024, False, 13421), (6, 1024, True, 8274),
                                                (48, 1024, False, 6.438e6), (48, 1024, True, 3.803e6)])
def test_flop_replay(oracle, n, k, nocov, per_step):
    """Per-step LAWN-41 counts of the reference's LAPACK/BLAS calls, as measured
    by interposing them on the reference build (F orthogonal, H=G=K=C=I)."""
    F = P.random_orthogonal(P.rng_for(1), n)
    pr = P.iid_problem(k, n, n, seed=5, F=F, K_kind="W", C_kind="W")
    oracle.lib.oracle_count_flops(1)
    oracle.oddeven(pr, (1 << 16) if nocov else 0)
    f = oracle.lib.oracle_flops() / k
    oracle.lib.oracle_count_flops(0)
    assert abs(f - per_step) / per_step < 1e-3


def test_checker_detects_planted_error(oracle):
    """A mistake planted in the oracle's output (one state component off by
    1e-7 relative) must fail the parity bar the GPU tests use."""
    d, meta = load("small_oddeven")
    tag = meta["cases"][-1]
    a = sub_arrays(d, tag)
    pr = P.problem_from_arrays(a, tag)
    xs, cs = oracle.oddeven(pr)
    assert max_rel_err(xs, list(a["out_state"])) < TOL_OE
    bad = [x.copy() for x in xs]
    j = len(bad) // 2
    bad[j][0] += 1e-7 * np.max(np.abs(bad[j]))
    assert max_rel_err(bad, list(a["out_state"])) > TOL_OE


def shape_outputs(a, pr):
    """Reference outputs of a shapes_oddeven case, split per step."""
    xs, cs, p, q = [], [], 0, 0
    for n in pr.n:
        xs.append(a["out_state_flat"][p:p + n])
        cs.append(a["out_cov_flat"][q:q + n * n].reshape(n, n, order="F"))
        p += n
        q += n * n
    return xs, cs


def test_shapes_fixture(oracle):
    """Per-step shapes of the reference's own catalogue (kalmanTest.m:80-95):
    rectangular H, changing state dimension, more observations than states.
    The oracle matches the reference where it ran (power-of-two lengths) and
    the global least-squares definition everywhere."""
    d, meta = load("shapes_oddeven")
    assert len(meta["reference_ok"]) >= 15
    kinds = set()
    for tag in meta["cases"]:
        a = sub_arrays(d, tag)
        pr = P.problem_from_arrays(a, tag)
        kinds |= {"rect_H" for h in pr.H if h is not None and h.shape[0] < h.shape[1]}
        kinds |= {"n_changes" for i in range(1, pr.k) if pr.n[i] != pr.n[i - 1]}
        kinds |= {"m>n" for i in range(pr.k) if pr.o[i] is not None and pr.o[i].shape[0] > pr.n[i]}
        xs, cs = oracle.oddeven(pr)
        lx, lc = uk_libs.dense_ls(pr)
        assert max(max_rel_err(xs, lx), max_rel_err(cs, lc)) < 1e-12, tag
        if tag in meta["reference_ok"]:
            rx, rc = shape_outputs(a, pr)
            assert max(max_rel_err(xs, rx), max_rel_err(cs, rc)) < TOL_OE, tag
    assert kinds == {"rect_H", "n_changes", "m>n"}


def pad_problem(pr):
    """The padded form libukhip's kalman_smooth_oddeven packs (every step N =
    max n_i components: [H_i 0], [F_i 0] and c_i in the first l_i of N rows,
    zero rows below; G_i as [G_i 0] plus N - n_i unit-weight rows z_i = 0;
    K_i, C_i extended block-diagonally with I), written here in numpy as an
    independent statement of that construction."""
    N = max(pr.n)
    H, F, c, K, Kt, G, o, Cc, Ct = ([None] * pr.k for _ in range(9))
    for i in range(pr.k):
        n = pr.n[i]
        if i > 0:
            l = pr.c[i].shape[0]
            H[i] = np.zeros((N, N)); H[i][:l, :n] = pr.H[i]
            F[i] = np.zeros((N, N)); F[i][:l, :pr.n[i - 1]] = pr.F[i]
            c[i] = np.zeros((N, 1)); c[i][:l] = pr.c[i]
            Kt[i] = pr.K_type[i]
            if Kt[i] == "w":
                K[i] = np.ones((N, 1)); K[i][:l] = pr.K[i]
            else:
                K[i] = np.eye(N); K[i][:l, :l] = pr.K[i]
        m = 0 if pr.o[i] is None else pr.o[i].shape[0]
        mp = m + N - n
        if mp == 0:
            continue
        G[i] = np.zeros((mp, N)); o[i] = np.zeros((mp, 1))
        if m:
            G[i][:m, :n] = pr.G[i]
            o[i][:m] = pr.o[i]
        for r in range(N - n):
            G[i][m + r, n + r] = 1.0
        Ct[i] = pr.C_type[i] if m else "w"
        if Ct[i] == "w":
            Cc[i] = np.ones((mp, 1))
            if m:
                Cc[i][:m] = pr.C[i]
        else:
            Cc[i] = np.eye(mp); Cc[i][:m, :m] = pr.C[i]
    return P.Problem(k=pr.k, n=[N] * pr.k, H=H, F=F, c=c, K=K, K_type=Kt, G=G, o=o, C=Cc, C_type=Ct,
                     name=pr.name + "_padded")


def test_padding_equivalence(oracle):
    """The padded problem's estimates and covariance blocks, restricted to each
    step's own n_i components, are the unpadded problem's (the padding
    components are decoupled and determined as zero)."""
    d, meta = load("shapes_oddeven")
    for tag in meta["cases"]:
        pr = P.problem_from_arrays(sub_arrays(d, tag), tag)
        xs, cs = oracle.oddeven(pr)
        px, pc = oracle.oddeven(pad_problem(pr))
        n = pr.n
        assert max_rel_err([px[i][:n[i]] for i in range(pr.k)], xs) < 1e-12, tag
        assert max_rel_err([pc[i][:n[i], :n[i]] for i in range(pr.k)], cs) < 1e-12, tag
        pad = [px[i][n[i]:] for i in range(pr.k) if n[i] < max(n)]
        assert all(np.max(np.abs(v)) < 1e-12 for v in pad), tag
