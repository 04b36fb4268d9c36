"""Algorithmic FLOP count of the odd-even smoother: a replay of the reference's
LAPACK/BLAS call shapes with the LAWN-41 operation counts.

The reference performs its block operations through dgeqrf, dormqr, dtrtrs,
dpotrf and dgemm (c/matrix_ops.c:360-597, c/kalman_base.c:96-105). This module
walks the same level recursion as c/kalman_oddeven_smoother.c:1358-1553 over
block *shapes* only (no arithmetic) and sums
    dgeqrf(M,N)      2MN^2 - 2/3 N^3   (M >= N), else 2NM^2 - 2/3 M^3
    dormqr(M,K,P)    P (4MK - 2K^2)
    dtrtrs(N,P)      P N^2
    dpotrf(N)        N^3 / 3
    dgemm(M,N,K)     2MNK
split into the whitening, the three elimination phases (phase 1 with the
single-step base case, phase 3 with the fix-last step of an odd level, as the
kernels k_phase1 and k_phase3 do them) and the back substitution. It is the definition of "algorithmic FLOPs" used for
the roofline: the GPU's own operation count (padding, redundant work) is not
credited.
"""


def _qr(M, N):
    if M == 0 or N == 0:
        return 0.0
    return 2.0 * M * N * N - 2.0 / 3.0 * N ** 3 if M >= N else 2.0 * N * M * M - 2.0 / 3.0 * M ** 3


def _ormqr(M, K, P):
    return P * (4.0 * M * K - 2.0 * K * K) if M > 0 else 0.0


def _trtrs(N, P):
    return float(P) * N * N


def _gemm(M, N, K):
    return 2.0 * M * N * K


def _weigh(kind, dim, cols):
    """kalman_covariance_matrix_weigh(cov (dim x dim), kind, A (dim x cols))."""
    if kind == "W":
        return _gemm(dim, cols, dim)
    if kind == "w":
        return 0.0
    if kind == "C":
        return dim ** 3 / 3.0 + _trtrs(dim, cols)
    return _trtrs(dim, cols)  # 'U', 'F'


def oddeven_flops(n, m_list, K_types=None, C_types=None, cov=True):
    """FLOPs of kalman_smooth_oddeven on one problem with constant state
    dimension n (H square), per-step observation counts m_list (0 =
    unobserved) and covariance types. Returns dict(weigh, phase1, phase2, phase3,
    eliminate = phase1 + phase2 + phase3, backsub, total)."""
    k = len(m_list)
    K_types = K_types or ["w"] * k
    C_types = C_types or ["w"] * k
    f = {"weigh": 0.0, "phase1": 0.0, "phase2": 0.0, "phase3": 0.0, "backsub": 0.0}
    # steps_weigh (c/kalman_oddeven_smoother.c:1620-1659)
    for i in range(k):
        if i > 0:
            f["weigh"] += 2 * _weigh(K_types[i], n, n) + _weigh(K_types[i], n, 1)
        if m_list[i] > 0:
            f["weigh"] += _weigh(C_types[i], m_list[i], n) + _weigh(C_types[i], m_list[i], 1)
    # per-step row counts of the chain: l (B, D, c rows), m (C, o rows)
    steps = [{"l": (n if i > 0 else 0), "m": m_list[i]} for i in range(k)]
    _level(steps, n, cov, f)
    f["eliminate"] = f["phase1"] + f["phase2"] + f["phase3"]
    f["total"] = f["weigh"] + f["eliminate"] + f["backsub"]
    return f


def _level(s, n, cov, f):
    L = len(s)
    E = "phase1"
    if L == 1:  # :1360-1423
        m = s[0]["m"]
        f[E] += _qr(m, n) + _ormqr(m, min(m, n), 1)
        f["backsub"] += _trtrs(n, 1) + 2 * _trtrs(n, n)
        return
    # phase 1 (:379-457): rows of R~, X, o (kept with step j) and D~, c (step j+1)
    for j in range(0, L, 2):
        si = s[j]
        if j == L - 1:
            m = si["m"]
            if m > 0:
                f[E] += _qr(m, n) + _ormqr(m, min(m, n), 1)
            si["rt"] = m
            si["o_rows"] = m
            continue
        sp = s[j + 1]
        rows = si["m"] + sp["l"]
        f[E] += _qr(rows, n)
        ll = min(n, rows)
        K = min(rows, n)
        f[E] += _ormqr(rows, K, 1) + _ormqr(rows, K, n)
        si["rt"] = ll
        si["o_rows"] = ll
        si["x_rows"] = ll
        sp["dt"] = rows - ll
        sp["c_rows"] = rows - ll
    # phase 2 (:460-523)
    E = "phase2"
    for j in range(0, L, 2):
        si = s[j]
        if j == 0:
            continue
        rows = si["l"] + si["rt"]
        f[E] += _qr(rows, n)
        mm = min(n, rows)
        K = min(rows, n)
        f[E] += _ormqr(rows, K, n) + _ormqr(rows, K, 1)
        si["bt_rows"] = mm
        si["z_rows"] = rows - mm
        si["o_rows2"] = rows - mm
        if j + 1 != L:
            f[E] += _ormqr(rows, K, n)
            si["y_rows"] = mm
            si["xt_rows"] = rows - mm
    # phase 3 (:526-560)
    E = "phase3"
    for j in range(0, L - 1, 2):
        sp = s[j + 1]
        rows = sp["dt"] + sp["m"]
        if rows == 0:
            sp["gt"] = 0
            sp["c_rows"] = 0
            continue
        f[E] += _qr(rows, n)
        ll = min(rows, n)
        f[E] += _ormqr(rows, min(rows, n), 1)
        sp["gt"] = ll
        sp["c_rows"] = ll
    # fix the last index (:1452-1476, D1 and D10 as in the oracle)
    if L % 2 == 1:
        sl, sm = s[L - 1], s[L - 2]
        rows = sm["gt"] + sl["z_rows"]
        f[E] += _qr(rows, n)
        keep = min(n, rows)
        f[E] += _ormqr(rows, min(rows, n), 1)
        split = min(max(n, sm["gt"]), keep)
        sm["gt"] = keep
        sm["c_rows"] = split
    # renaming (:563-597) and recursion
    sub = []
    for j in range(0, L - 1, 2):
        sp = s[j + 1]
        ns = {"m": sp["gt"], "l": 0}
        if j != 0:
            ns["l"] = s[j]["z_rows"]
        sub.append(ns)
    _level(sub, n, cov, f)
    # back substitution and covariances (:614-975)
    B = "backsub"
    for j in range(0, L, 2):
        si = s[j]
        if j == 0:
            xr = si["x_rows"]
            f[B] += _gemm(xr, 1, n) + _trtrs(n, 1)
            if cov:
                f[B] += _gemm(xr, n, n) + _gemm(xr, xr, n) + 2 * _trtrs(n, n) + _gemm(xr, n, n) + _trtrs(n, n)
        elif j != L - 1:
            b, y = si["bt_rows"], si["y_rows"]
            f[B] += _gemm(b, 1, n) + _gemm(y, 1, n) + _trtrs(n, 1)
            if cov:
                f[B] += (_gemm(b, n, n) + _gemm(b, b, n) + _gemm(y, n, n) + _gemm(y, y, n) + _gemm(b, n, n)
                         + _gemm(b, y, n) + 2 * _trtrs(n, n) + _gemm(b, n, n) + _gemm(y, n, n) + _trtrs(n, n)
                         + _gemm(y, n, n) + _gemm(b, n, n) + _trtrs(n, n))
        else:
            b = si["bt_rows"]
            f[B] += _gemm(b, 1, n) + _trtrs(n, 1)
            if cov:
                f[B] += _gemm(b, n, n) + _gemm(b, b, n) + 2 * _trtrs(n, n) + _gemm(b, n, n) + _trtrs(n, n)


def oddeven_bytes(n, m_list):
    """Algorithmic HBM bytes of one smoothing in FP64: per-step inputs F (n^2),
    c (n), G (m n), o (m) and outputs state (n) and covariance (n^2)."""
    k = len(m_list)
    inputs = sum((n * n + n if i > 0 else 0) + m * n + m for i, m in enumerate(m_list))
    return 8.0 * (inputs + k * (n + n * n))


def oddeven_backsub_bytes(n, k, cov=True):
    """Algorithmic HBM bytes of the back substitution (all levels) of one
    problem of k steps in FP64: per even step j of a level, it reads R_j, cf_j,
    B~_j (j > 0) and Y_j (j < L-1), the neighbours' estimates, covariances and
    (both neighbours present) their cross-covariance; it writes x_j, S_j and
    the cross-covariances with its neighbours. Odd steps are written by the
    level above, so each output is counted once."""
    nn = n * n
    total = 0
    L = k
    while L >= 1:
        for j in range(0, L, 2):
            nb = (j > 0) + (j < L - 1)
            r = nn + n + nb * nn + nb * n
            w = n
            if cov:
                r += nb * nn + (nn if nb == 2 else 0)
                w += nn + nb * nn
            total += r + w
        if L == 1:
            break
        L //= 2
    return 8.0 * total


def oddeven_phase2_bytes(n, k, fp32_levels=0, level0_diag_synth=False):
    """Algorithmic HBM bytes of the odd-even phase 2 (QR of [D_j; R~_j], the
    reference's c/kalman_oddeven_smoother.c:460-523) over all levels of one
    problem of k steps: per even step j > 0 of a level it reads D_j, R~_j,
    X_j, B_j (n x n each), c_j and o_j, and writes R, B~, Y, Z_j, X~_j (n x n
    each; Z_j and o_j of a lone last step go to a scratch of the same size)
    and cf_j, c~_j. Levels below fp32_levels hold 4-byte elements, the others
    8. level0_diag_synth: at level 0, D_j is stored as its diagonal and B_j is
    formed from the FP64 K_j scaling (n values; F, one block per problem, is
    not counted)."""
    nn = n * n
    total = 0.0
    L, lev = k, 0
    while L >= 2:
        e = 4.0 if lev < fp32_levels else 8.0
        if lev == 0 and level0_diag_synth:
            r = e * (n + 2 * nn + 2 * n) + 8.0 * n
        else:
            r = e * (4 * nn + 2 * n)
        total += (L - 1) // 2 * (r + e * (5 * nn + 2 * n))  # even j = 2, 4, ... < L
        L //= 2
        lev += 1
    return total


def _mldivide(N, P):
    """matrix_create_mldivide(A (N x N), B (N x P)) of the reference
    (c/matrix_ops.c:610-630): dgeqrf, dormqr, dtrtrs."""
    return _qr(N, N) + _ormqr(N, N, P) + _trtrs(N, P)


def associative_flops(n, m_list):
    """FLOPs of kalman_smooth_associative on one problem (constant n, square
    H, 'w' covariances so that the explicit covariances cost nothing), per
    kernel class of the GPU path, from the reference's matrix operations:
      felem  build_filtering_element_new, observed steps
             (c/kalman_associative_smoother.c:370-505), per step with m rows;
      fscan  filteringAssociativeOperation (:889-980) per combine, two
             combines per element (the work-efficient parallel scan the
             reference runs through prefix_sums_pointers, c/parallel_tbb.cpp:78-121;
             a sequential scan would do one);
      selem  build_smoothing_element_new (:697-753), every step but the last;
      sscan  smoothingAssociativeOperation (:982-1016), two per element;
      prior  the first step's information filter (:335-368), once.
    weigh, model, filtered and results are the GPU path's own data movement
    (no reference arithmetic): zero. Vector terms are included."""
    k = len(m_list)
    f = dict(weigh=0.0, model=0.0, prior=0.0, felem=0.0, fscan=0.0, filtered=0.0, selem=0.0, sscan=0.0,
             results=0.0)
    g = _gemm
    for i, m in enumerate(m_list[1:], start=1):
        if m == 0:
            continue
        fe = (g(n, m, n) + g(m, m, n) + _mldivide(m, n) + g(n, m, n)   # K_i G^T, G K G^T, S^-T G, K
              + g(m, n, n) + g(n, n, m) + g(m, 1, n) + g(n, 1, m)      # G F, K G F, G c, K (o - G c)
              + g(n, n, m) + g(n, n, n) + g(m, 1, n) + g(n, m, n)      # K G, K G K_i, G c, F^T G^T S^-1
              + g(n, 1, m) + g(m, n, n) + g(n, n, m))                  # e, G F, J
        f["felem"] += fe
    combine_f = (g(n, n, n) + _mldivide(n, n) + g(n, n, n) + _mldivide(n, n) + g(n, n, n) + g(n, 1, n)
                 + g(n, 1, n) + g(n, n, n) + g(n, n, n) + g(n, 1, n) + g(n, 1, n) + g(n, n, n) + g(n, n, n))
    combine_s = g(n, n, n) + g(n, 1, n) + g(n, n, n) + g(n, n, n)
    f["fscan"] = 2.0 * max(k - 1, 0) * combine_f
    f["sscan"] = 2.0 * k * combine_s
    selem = g(n, n, n) + g(n, n, n) + _mldivide(n, n) + g(n, 1, n) + g(n, 1, n) + g(n, n, n) + g(n, n, n)
    f["selem"] = max(k - 1, 0) * selem
    m0 = m_list[0] if m_list else 0
    f["prior"] = g(n, m0, n) + g(n, n, m0) + _mldivide(n, n) if m0 else 0.0
    f["total"] = sum(v for key, v in f.items())
    return f
