"""Seeded synthetic smoothing problems.

A Problem is a list of step equations in the reference's form
(c/kalman.h:56-75):  H_i u_i = F_i u_{i-1} + c_i + eps_i  (i > 0, Cov eps_i from
(K_i, K_type_i)),  o_i = G_i u_i + delta_i  (Cov delta_i from (C_i, C_type_i);
o_i is None on unobserved steps).

Generators for uniform problems also keep the per-step matrices packed in
column-major-per-step arrays (`packed`), which the device API uploads as is.
Random streams come from numpy's Philox bit generator, so a (kind, seed)
pair names the same inputs on every machine with this image.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Problem:
    k: int
    n: list
    H: list
    F: list
    c: list
    K: list
    K_type: list
    G: list
    o: list
    C: list
    C_type: list
    truth: np.ndarray = None
    packed: dict = field(default_factory=dict)
    name: str = ""

    @property
    def steps(self):
        return self.k


def rng_for(seed):
    return np.random.Generator(np.random.Philox(seed))


def random_orthogonal(rng, n):
    """Q of a Householder QR of an N(0,1) matrix, times diag(sign(diag R))."""
    a = rng.standard_normal((n, n))
    q, r = np.linalg.qr(a)
    s = np.sign(np.diag(r))
    s[s == 0] = 1.0
    return q * s


def _cov_repr(kind, var, dim):
    """Covariance var*I of size dim in representation `kind`."""
    if kind == "w":
        return np.full((dim, 1), 1.0 / np.sqrt(var)), "w"
    if kind == "W":
        return np.eye(dim) / np.sqrt(var), "W"
    if kind == "C":
        return np.eye(dim) * var, "C"
    if kind in ("U", "F"):
        return np.eye(dim) * np.sqrt(var), kind
    raise ValueError(kind)


def _sqrt_cov(cov, kind):
    """A matrix S with S S^T = the covariance represented by (cov, kind)."""
    if kind == "w":
        return np.diag(1.0 / cov[:, 0])
    if kind == "W":
        return np.linalg.inv(cov)
    if kind == "C":
        return np.linalg.cholesky(cov)
    return cov  # 'U'/'F': lower factor


def uniform_problem(k, n, m, *, seed, F=None, F_fn=None, G_fn=None, Kvar=1.0, Cvar=1.0,
                    K_kind="w", C_kind="w", unobserved=None, H_identity=True, c_zero=True,
                    name=""):
    """A problem with constant n and m and scalar*I noise covariances.

    F_fn(rng, i) / G_fn(rng, i) give per-step matrices (default: the fixed F,
    G = I). `unobserved(i)` marks steps without observations.
    """
    rng = rng_for(seed)
    Fp = np.zeros((k, n, n))  # column-major per step: Fp[i] = F_i^T in C order
    Gp = np.zeros((k, n, m))
    cp = np.zeros((k, n))
    op = np.zeros((k, m))
    obs = np.ones(k, dtype=bool)
    Kmat, K_type = _cov_repr(K_kind, Kvar, n)
    Cmat, C_type = _cov_repr(C_kind, Cvar, m)
    Ks = _sqrt_cov(Kmat, K_type)
    Cs = _sqrt_cov(Cmat, C_type)
    u = rng.standard_normal(n)
    truth = np.zeros((k, n))
    Hm = np.eye(n)
    for i in range(k):
        if i > 0:
            Fi = F if F_fn is None else F_fn(rng, i)
            ci = np.zeros(n) if c_zero else rng.standard_normal(n)
            u = Fi @ u + ci + Ks @ rng.standard_normal(n)
            Fp[i] = Fi.T
            cp[i] = ci
        truth[i] = u
        if unobserved is not None and unobserved(i):
            obs[i] = False
            continue
        Gi = np.eye(m, n) if G_fn is None else G_fn(rng, i)
        Gp[i] = Gi.T
        op[i] = Gi @ u + Cs @ rng.standard_normal(m)
    P = Problem(k=k, n=[n] * k,
                H=[None] + [Hm] * (k - 1),
                F=[None] + [Fp[i].T for i in range(1, k)],
                c=[None] + [cp[i].reshape(n, 1) for i in range(1, k)],
                K=[None] + [Kmat] * (k - 1),
                K_type=[None] + [K_type] * (k - 1),
                G=[Gp[i].T if obs[i] else None for i in range(k)],
                o=[op[i].reshape(m, 1) if obs[i] else None for i in range(k)],
                C=[Cmat if obs[i] else None for i in range(k)],
                C_type=[C_type if obs[i] else None for i in range(k)],
                truth=truth, name=name)
    P.packed = dict(n=n, m=m, F=Fp, G=Gp, c=cp, o=op, observed=obs,
                    K=Kmat, K_type=K_type, C=Cmat, C_type=C_type)
    return P


# ---- the benchmark configurations --------------------------------------------------

def config_rotation6(k=100_000, seed=1):
    """n = m = 6, fixed random orthogonal F, H = G = K = C = I."""
    F = random_orthogonal(rng_for(seed + 1000), 6)
    return uniform_problem(k, 6, 6, seed=seed, F=F, K_kind="W", C_kind="W",
                           name=f"n6_orthogonal_k{k}")


def config_assoc48(k=200_000, seed=2):
    """n = 48, m = 24, F = 0.99 Q, G ~ N(0, 1/48) per step, K = I, C = 0.1 I,
    every 10th step (i mod 10 == 9) unobserved."""
    F = 0.99 * random_orthogonal(rng_for(seed + 1000), 48)
    return uniform_problem(k, 48, 24, seed=seed, F=F,
                           G_fn=lambda r, i: r.standard_normal((24, 48)) / np.sqrt(48.0),
                           Kvar=1.0, Cvar=0.1, unobserved=lambda i: i % 10 == 9,
                           name=f"n48_m24_k{k}")


def config_precision64(k=200_000, seed=4):
    """n = 64, m = 32, F = Q diag(logspace(0,-3,64)), G ~ N(0, 1/64), K = 1e-2 I, C = I."""
    F = random_orthogonal(rng_for(seed + 1000), 64) @ np.diag(np.logspace(0, -3, 64))
    return uniform_problem(k, 64, 32, seed=seed, F=F,
                           G_fn=lambda r, i: r.standard_normal((32, 64)) / 8.0,
                           Kvar=1e-2, Cvar=1.0, name=f"n64_m32_k{k}")


def batch_problems(count=4096, seed=3, lmin=64, lmax=16384):
    """Independent problems: n uniform over {4,8,...,32}, m = n/2, lengths
    log-uniform in [lmin, lmax], F = s Q with s ~ U[0.95, 1], G ~ N(0, 1/n)."""
    rng = rng_for(seed)
    probs = []
    for p in range(count):
        n = int(rng.choice(np.arange(4, 33, 4)))
        m = n // 2
        L = int(np.exp(rng.uniform(np.log(lmin), np.log(lmax))))
        F = rng.uniform(0.95, 1.0) * random_orthogonal(rng, n)
        probs.append(uniform_problem(
            L, n, m, seed=int(rng.integers(1 << 62)), F=F,
            G_fn=(lambda nn, mm: (lambda r, i: r.standard_normal((mm, nn)) / np.sqrt(nn)))(n, m),
            name=f"batch{p}_n{n}_k{L}"))
    return probs


# ---- the reference's shape catalogue (matlab/examples/kalmanTest.m:80-95) ----------

def random_problem(seed, k, n, m, *, K_kind="W", C_kind="W", unobserved=(), H_kind="I",
                   m_of=None, scale=1.0):
    """General small problem with random dense covariance factors.

    H_kind: 'I' identity, 'R' random well-conditioned square. m_of(i) gives a
    per-step observation count (default m); steps in `unobserved` have none.
    """
    rng = rng_for(seed)
    H, F, c, K, Kt, G, o, Cc, Ct = ([None] * k for _ in range(9))
    truth = np.zeros((k, n))
    u = rng.standard_normal(n)

    def cov_of(kind, dim):
        a = rng.standard_normal((dim, dim)) * 0.3 + np.eye(dim) * 1.5
        cov = a @ a.T * scale
        if kind == "C":
            return cov, "C", np.linalg.cholesky(cov)
        if kind == "W":
            Lc = np.linalg.cholesky(cov)
            return np.linalg.inv(Lc), "W", Lc
        if kind == "w":
            d = rng.uniform(0.5, 2.0, size=(dim, 1))
            return d, "w", np.diag(1.0 / d[:, 0])
        Lc = np.linalg.cholesky(cov)
        return Lc, kind, Lc

    for i in range(k):
        if i > 0:
            Fi = rng.standard_normal((n, n)) / np.sqrt(n)
            Hi = np.eye(n) if H_kind == "I" else np.eye(n) + 0.2 * rng.standard_normal((n, n))
            ci = rng.standard_normal(n) * 0.1
            Ki, kt, Ks = cov_of(K_kind, n)
            u = np.linalg.solve(Hi, Fi @ u + ci + Ks @ rng.standard_normal(n))
            H[i], F[i], c[i], K[i], Kt[i] = Hi, Fi, ci.reshape(n, 1), Ki, kt
        truth[i] = u
        mi = m if m_of is None else m_of(i)
        if i in unobserved or mi == 0:
            continue
        Gi = rng.standard_normal((mi, n))
        Ci, ct, Cs = cov_of(C_kind, mi)
        G[i], o[i], Cc[i], Ct[i] = Gi, (Gi @ u + Cs @ rng.standard_normal(mi)).reshape(mi, 1), Ci, ct
    return Problem(k=k, n=[n] * k, H=H, F=F, c=c, K=K, K_type=Kt, G=G, o=o, C=Cc, C_type=Ct,
                   truth=truth, name=f"rand_s{seed}_k{k}_n{n}_m{m}")


# ---- deterministic large problems and (de)serialization -----------------------------

def iid_problem(k, n, m, *, seed, F, G_scale=None, Kvar=1.0, Cvar=1.0, K_kind="w", C_kind="w",
                unobserved=None, name="", lists=True):
    """Like uniform_problem, but observations are i.i.d. N(0,1) vectors and G_i
    is either I (G_scale None) or N(0,1)*G_scale per step. Only Philox normals
    and exact IEEE scalings are used, so the inputs are bit-identical on every
    machine with this numpy (no BLAS on the generation path). F is the same on
    every step (packed["F_shared"] holds it once). lists=False keeps only the
    packed arrays (for the device batch API on large inputs)."""
    rng = rng_for(seed)
    Fp = np.broadcast_to(np.asfortranarray(F).T, (k, n, n))
    if G_scale is None:
        Gp = np.broadcast_to(np.eye(m, n).T, (k, n, m))
    else:
        Gp = np.swapaxes(rng.standard_normal((k, m, n)) * G_scale, 1, 2).copy()
    op = rng.standard_normal((k, m))
    cp = np.zeros((k, n))
    obs = np.ones(k, dtype=bool)
    if unobserved is not None:
        obs = np.array([not unobserved(i) for i in range(k)])
    Kmat, K_type = _cov_repr(K_kind, Kvar, n)
    Cmat, C_type = _cov_repr(C_kind, Cvar, m)
    packed = dict(n=n, m=m, F=Fp, G=Gp, c=cp, o=op, observed=obs,
                  K=Kmat, K_type=K_type, C=Cmat, C_type=C_type, F_shared=np.asfortranarray(F))
    if not lists:
        P = Problem(k=k, n=[n] * k, H=[], F=[], c=[], K=[], K_type=[], G=[], o=[], C=[], C_type=[], name=name)
        P.packed = packed
        return P
    P = Problem(k=k, n=[n] * k,
                H=[None] + [np.eye(n)] * (k - 1),
                F=[None] + [F] * (k - 1),
                c=[None] + [cp[i].reshape(n, 1) for i in range(1, k)],
                K=[None] + [Kmat] * (k - 1),
                K_type=[None] + [K_type] * (k - 1),
                G=[Gp[i].T if obs[i] else None for i in range(k)],
                o=[op[i].reshape(m, 1) if obs[i] else None for i in range(k)],
                C=[Cmat if obs[i] else None for i in range(k)],
                C_type=[C_type if obs[i] else None for i in range(k)],
                name=name)
    P.packed = packed
    return P


def batch_specs(count=4096, seed=3, lmin=64, lmax=16384, nmin=4, nmax=32, nstep=4):
    """The batching workload's problem list, without the data: (n, m, k, F
    scale, seed) per problem. n uniform over {nmin, nmin+nstep, ..., nmax},
    m = n/2, k log-uniform in [lmin, lmax]."""
    rng = rng_for(seed)
    specs = []
    for p in range(count):
        n = int(rng.choice(np.arange(nmin, nmax + 1, nstep)))
        k = int(np.exp(rng.uniform(np.log(lmin), np.log(lmax))))
        specs.append(dict(n=n, m=max(n // 2, 1), k=k, s=float(rng.uniform(0.95, 1.0)),
                          seed=int(rng.integers(1 << 62)), name=f"batch{p}_n{n}_k{k}"))
    return specs


def batch_problem(spec, lists=True):
    """One problem of batch_specs: F = s Q (Q random orthogonal), G ~ N(0, 1/n)
    per step, K = C = I, i.i.d. N(0,1) observations."""
    n, m = spec["n"], spec["m"]
    F = spec["s"] * random_orthogonal(rng_for(spec["seed"] ^ 0x5A5A), n)
    return iid_problem(spec["k"], n, m, seed=spec["seed"], F=F, G_scale=1.0 / np.sqrt(n),
                       name=spec["name"], lists=lists)


def serve_specs(count, seed=5, lmin=100, lmax=50_000, nmin=3, nmax=16, unobserved=0.05):
    """The serving workload's request list, without the data: n uniform over
    {nmin..nmax}, m uniform over [1, n], length log-uniform in [lmin, lmax],
    each step unobserved with probability `unobserved`."""
    rng = rng_for(seed)
    specs = []
    for p in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        m = int(rng.integers(1, n + 1))
        k = int(np.exp(rng.uniform(np.log(lmin), np.log(lmax))))
        specs.append(dict(n=n, m=m, k=k, seed=int(rng.integers(1 << 62)), unobserved=unobserved,
                          name=f"req{p}_n{n}_m{m}_k{k}"))
    return specs


def serve_problem(spec, lists=True):
    """One request of serve_specs: F = 0.99 Q_n, G ~ N(0, 1/n) per step, K = I,
    C = 0.1 I, i.i.d. N(0,1) observations; the first steps are observed until
    they determine the state (so both smoothers accept the request)."""
    n, m, k = spec["n"], spec["m"], spec["k"]
    rng = rng_for(spec["seed"] ^ 0x3C3C)
    unob = rng.random(k) < spec["unobserved"]
    unob[: (n + m - 1) // m] = False
    F = 0.99 * random_orthogonal(rng, n)
    return iid_problem(k, n, m, seed=spec["seed"], F=F, G_scale=1.0 / np.sqrt(n), Kvar=1.0, Cvar=0.1,
                       unobserved=lambda i: bool(unob[i]), name=spec["name"], lists=lists)


def input_digest(problem):
    """SHA-256 over every input matrix of every step, in step order."""
    import hashlib
    h = hashlib.sha256()
    for i in range(problem.k):
        for name in ("H", "F", "c", "K", "G", "o", "C"):
            a = getattr(problem, name)[i]
            if a is None:
                h.update(b"-")
            else:
                a = np.asarray(a, dtype=np.float64)
                h.update(np.asarray(a.shape, dtype=np.int64).tobytes())
                h.update(np.ascontiguousarray(a.T).tobytes())
        for name in ("K_type", "C_type"):
            t = getattr(problem, name)[i]
            h.update((t or "-").encode())
    return h.hexdigest()


_FIELDS = ("H", "F", "c", "K", "G", "o", "C")


def problem_to_arrays(problem):
    """Flat arrays (for .npz fixtures) holding every input exactly."""
    shapes, data, types = [], [], []
    for i in range(problem.k):
        for f in _FIELDS:
            a = getattr(problem, f)[i]
            if a is None:
                shapes.append((-1, -1))
            else:
                a = np.asarray(a, dtype=np.float64).reshape(np.shape(a)[0], -1)
                shapes.append(a.shape)
                data.append(np.ascontiguousarray(a.T).ravel())
        types.append([ord(problem.K_type[i] or "-"), ord(problem.C_type[i] or "-")])
    return dict(in_n=np.asarray(problem.n, dtype=np.int32),
                in_shapes=np.asarray(shapes, dtype=np.int32),
                in_data=np.concatenate(data) if data else np.zeros(0),
                in_types=np.asarray(types, dtype=np.int32))


def problem_from_arrays(d, name=""):
    n = [int(x) for x in d["in_n"]]
    k = len(n)
    shapes = d["in_shapes"].reshape(k, len(_FIELDS), 2)
    data = d["in_data"]
    pos = 0
    cols = {f: [None] * k for f in _FIELDS}
    for i in range(k):
        for j, f in enumerate(_FIELDS):
            r, c = shapes[i, j]
            if r < 0:
                continue
            a = data[pos:pos + r * c].reshape(c, r).T
            pos += r * c
            cols[f][i] = np.asfortranarray(a)
    t = d["in_types"]
    kt = [None if x == ord("-") else chr(x) for x in t[:, 0]]
    ct = [None if x == ord("-") else chr(x) for x in t[:, 1]]
    return Problem(k=k, n=n, K_type=kt, C_type=ct, name=name, **cols)


def shape_problem(seed, rows, *, K_kind="W", C_kind="W"):
    """A problem with per-step shapes, as matlab/examples/compare.m:30-71 builds
    them from a kalmanTest.m row table: step i has state dimension n_i, an
    evolution equation of l_i rows (H_i = eye(l_i, n_i), F_i the first l_i x
    n_{i-1} block of a random orthogonal matrix, c_i ~ N(0,1)) and m_i
    observation rows (G_i, o_i ~ N(0,1)); `rows` lists (n_i, l_i, m_i).
    Covariances are random SPD matrices in representation K_kind / C_kind."""
    rng = rng_for(seed)
    k = len(rows)
    H, F, c, K, Kt, G, o, Cc, Ct = ([None] * k for _ in range(9))

    def cov_of(kind, dim):
        a = rng.standard_normal((dim, dim)) * 0.3 + np.eye(dim) * 1.5
        cov = a @ a.T
        if kind == "C":
            return cov, "C"
        if kind == "W":
            return np.linalg.inv(np.linalg.cholesky(cov)), "W"
        if kind == "w":
            return rng.uniform(0.5, 2.0, size=(dim, 1)), "w"
        return np.linalg.cholesky(cov), kind

    for i, (n, l, m) in enumerate(rows):
        if i > 0:
            npv = rows[i - 1][0]
            d = max(l, npv)
            q, r = np.linalg.qr(rng.standard_normal((d, d)))
            H[i] = np.eye(l, n)
            F[i] = np.asfortranarray(q[:l, :npv])
            c[i] = rng.standard_normal((l, 1))
            K[i], Kt[i] = cov_of(K_kind, l)
        if m > 0:
            G[i] = rng.standard_normal((m, n))
            o[i] = rng.standard_normal((m, 1))
            Cc[i], Ct[i] = cov_of(C_kind, m)
    return Problem(k=k, n=[r[0] for r in rows], H=H, F=F, c=c, K=K, K_type=Kt, G=G, o=o, C=Cc, C_type=Ct,
                   name=f"shape_s{seed}_k{k}")


def shape_catalogue():
    """Row tables (n_i, l_i, m_i) of matlab/examples/kalmanTest.m:80-95, plus
    power-of-two-length variants (the reference's odd-even smoother runs only
    on power-of-two lengths, SURVEY defect D1) and larger random tables with
    rectangular H, changing dimensions and more observations than states."""
    def uniform(n, l, m, k):
        return [(n, l, m)] * k
    cat = {
        "u1117": uniform(1, 1, 1, 7), "u2221": uniform(2, 2, 2, 1), "u2227": uniform(2, 2, 2, 7),
        "u3331": uniform(3, 3, 3, 1), "u3334": uniform(3, 3, 3, 4), "u3335": uniform(3, 3, 3, 5),
        "u3313": uniform(3, 3, 1, 3), "u4422": uniform(4, 4, 2, 2),
        "u2129": uniform(2, 1, 2, 9), "u2128": uniform(2, 1, 2, 8),
        "u3239": uniform(3, 2, 3, 9), "u3238": uniform(3, 2, 3, 8),
        "u3139": uniform(3, 1, 3, 9), "u31316": uniform(3, 1, 3, 16),
        "mid3obs": [(3, 0, 2), (3, 1, 2), (3, 1, 3), (3, 1, 2)],
        "grow2": [(1, 1, 0), (1, 1, 2)],
        "grow7": [(1, 1, 0), (1, 1, 1), (2, 1, 2), (3, 2, 3), (3, 3, 0), (3, 2, 3), (4, 2, 3)],
        "grow8": [(1, 1, 0), (1, 1, 1), (2, 1, 2), (3, 2, 3), (3, 3, 0), (3, 2, 3), (4, 2, 3), (4, 3, 4)],
        "wide8": uniform(2, 2, 4, 8), "wide16": uniform(3, 3, 5, 16),
    }
    rng = rng_for(777)
    for t in range(4):
        # larger tables: n_i in 4..12, l_i in [n_i - 3, n_i], m_i in [n_i - l_i, n_i + 3] (step 0: l_0 = 0), so
        # that [H_i; G_i] has full column rank and every level-0 block the elimination inverts is nonsingular
        k = (32, 64, 64, 128)[t]
        rows = []
        for i in range(k):
            n = int(rng.integers(4, 13))
            l = int(rng.integers(max(1, n - 3), n + 1))
            m = int(rng.integers(n if i == 0 else n - l, n + 4))
            rows.append((n, l, m))
        cat[f"rand{t}_k{k}"] = rows
    return cat


def packed_from_lists(pr):
    """The packed (batch API) form of a list-form Problem with one state
    dimension and one covariance representation per kind: per-step F, H, c,
    K, G, o, C arrays (column-major per step), observation rows m = max m_i
    with the per-step counts in m_i. Returns pr with `packed` set."""
    n, k = pr.n[0], pr.k
    assert all(x == n for x in pr.n)
    ms = [0 if o is None else o.shape[0] for o in pr.o]
    m = max(ms)
    kt = next((t for t in pr.K_type if t), "w")
    ct = next((t for t in pr.C_type if t), "w")
    assert all(t in (None, kt) for t in pr.K_type) and all(t in (None, ct) for t in pr.C_type)
    F = np.zeros((k, n, n))
    H = np.zeros((k, n, n))
    c = np.zeros((k, n))
    K = np.zeros((k, n)) if kt == "w" else np.zeros((k, n, n))
    G = np.zeros((k, n, max(m, 1)))
    o = np.zeros((k, max(m, 1)))
    C = np.zeros((k, max(m, 1))) if ct == "w" else np.zeros((k, max(m, 1), max(m, 1)))
    H_ident = True
    for i in range(k):
        if i > 0:
            F[i] = np.asarray(pr.F[i]).T
            H[i] = np.asarray(pr.H[i]).T
            H_ident &= bool(np.array_equal(pr.H[i], np.eye(n)))
            c[i] = np.ravel(pr.c[i])
            K[i] = np.ravel(pr.K[i]) if kt == "w" else np.asarray(pr.K[i]).T
        else:
            H[i] = np.eye(n)
            K[i] = 1.0 if kt == "w" else np.eye(n)
        mi = ms[i]
        if mi:
            G[i, :, :mi] = np.asarray(pr.G[i]).T
            o[i, :mi] = np.ravel(pr.o[i])
            if ct == "w":
                C[i, :mi] = np.ravel(pr.C[i])
            else:
                C[i, :mi, :mi] = np.asarray(pr.C[i]).T
    pr.packed = dict(n=n, m=m, F=F, H=None if H_ident else H, c=c, o=o, G=G, observed=np.array(ms) > 0,
                     m_i=np.array(ms, dtype=np.int32), K_steps=K, K_type=kt, C_steps=C, C_type=ct)
    return pr

