"""CPU check of the tridiagonal eigensolver design used on the GPU (eigensolver
mode 3, rsdft_amd/csrc/tridiag.hip): a numpy restatement of the same steps -
multisection on dstebz Sturm counts with its pivmin guard, inverse iteration
with dgttrf/dgttrs and pivots raised to eps |T|, modified Gram-Schmidt inside
clusters closer than 1e-7 |T|, one Newton-Schulz step - against
numpy.linalg.eigh on tridiagonal matrices that stress it: random spectra,
triplets 1e-9 apart and exactly repeated uncoupled blocks.  Bounds as in the
GPU eigensolver tests: eigenvalues 1e-12 |T|, residual 1e-11 |T|,
orthogonality 1e-12.  This pins the algorithm; the GPU kernels themselves are
checked by tests/test_gpu_eigh.py.
"""
import numpy as np
import pytest

EPS = 2.0 ** -53
SAFMIN = 2.2250738585072014e-308
SHIFTS = 128          # 64 lanes x 2 shifts per round
ROUNDS = 2            # inverse-iteration solves
CLUSTER = 1e-7        # relative gap below which vectors are orthogonalised


def bisect(d, e):
    n = len(d)
    el = np.r_[0.0, np.abs(e)]
    er = np.r_[np.abs(e), 0.0]
    gl, gu = np.min(d - el - er), np.max(d + el + er)
    e2 = er ** 2
    tn = max(abs(gl), abs(gu))
    piv = SAFMIN * max(1.0, e2.max() if n > 1 else 0.0)
    lo = np.full(n, gl - 2 * EPS * tn - 2 * piv)
    hi = np.full(n, gu + 2 * EPS * tn + 2 * piv)
    j = np.arange(n)
    for _ in range(64):
        act = ~(hi - lo <= 2 * EPS * np.maximum(abs(lo), abs(hi)) + 0.5 * EPS * tn)
        if not act.any():
            break
        h = (hi - lo) / (SHIFTS + 1)
        sg = lo[:, None] + (np.arange(SHIFTS)[None, :] + 1) * h[:, None]
        q = d[0] - sg
        q = np.where(abs(q) < piv, -piv, q)
        cnt = (q < 0).astype(int)
        for i in range(1, n):
            q = (d[i] - sg) - e2[i - 1] / q
            q = np.where(abs(q) < piv, -piv, q)
            cnt += q < 0
        ok = cnt >= (j[:, None] + 1)
        kl = np.where(ok.any(1), ok.argmax(1), SHIFTS)
        nlo = lo + kl * h
        nhi = np.where(kl < SHIFTS, lo + (kl + 1) * h, hi)
        lo, hi = np.where(act, nlo, lo), np.where(act, nhi, hi)
    return 0.5 * (lo + hi), tn


def inverse_iteration(d, e, w, tn, X, first):
    """One dgttrs solve per eigenvector (columns), LU of T - w_j I by dgttrf."""
    n, m = len(d), len(w)
    ee = np.r_[e, 0.0]
    ptol = max(EPS * tn, SAFMIN)
    dc = d[0] - w
    duc = np.full(m, ee[0])
    FL = np.zeros((n, m)); FD = np.zeros((n, m)); FU = np.zeros((n, m))
    FU2 = np.zeros((n, m)); FP = np.zeros((n, m), bool)
    for i in range(n - 1):
        li, dn, dun = ee[i], d[i + 1] - w, ee[i + 1] if i < n - 2 else 0.0
        keep = np.abs(dc) >= abs(li)
        num = np.where(keep, li, dc)
        den = np.where(keep, dc, li)
        f = num / np.where(den == 0, 1.0, den)
        FL[i], FD[i] = f, np.where(keep, dc, li)
        FU[i], FU2[i], FP[i] = np.where(keep, duc, dn), np.where(keep, 0.0, dun), ~keep
        dc, duc = np.where(keep, dn - f * duc, duc - f * dn), np.where(keep, dun, -f * dun)
    FD[n - 1] = dc
    FD = np.where(np.abs(FD) < ptol, np.where(FD < 0, -ptol, ptol), FD)
    B = X.copy()
    for i in range(n - 1):
        bi, bn = B[i].copy(), B[i + 1].copy()
        B[i] = np.where(FP[i], bn, bi)
        B[i + 1] = np.where(FP[i], bi, bn) - FL[i] * B[i]
    x1 = np.zeros(m); x2 = np.zeros(m)
    for i in range(n - 1, -1, -1):
        x = (B[i] - (FU[i] * x1 if i < n - 1 else 0.0) - (FU2[i] * x2 if i < n - 2 else 0.0)) / FD[i]
        B[i], x2, x1 = x, x1, x
    return B / np.linalg.norm(B, axis=0)


def cluster_mgs(w, tn, Z):
    n, j = len(w), 0
    while j < n:
        m = j
        while m + 1 < n and w[m + 1] - w[m] < CLUSTER * tn:
            m += 1
        for k in range(j + 1, m + 1):
            for i0 in range(j, k):
                Z[:, k] -= (Z[:, i0] @ Z[:, k]) * Z[:, i0]
            Z[:, k] /= np.linalg.norm(Z[:, k])
        j = m + 1
    return Z


def tri_eig(d, e):
    w, tn = bisect(d, e)
    rng = np.random.default_rng(1)
    Z = rng.uniform(-1.0, 1.0, (len(d), len(d)))
    for r in range(ROUNDS):
        Z = cluster_mgs(w, tn, inverse_iteration(d, e, w, tn, Z, r == 0))
    return w, Z @ (1.5 * np.eye(len(d)) - 0.5 * Z.T @ Z)


def check(d, e):
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    wr = np.linalg.eigvalsh(T)
    scale = np.max(np.abs(wr))
    w, V = tri_eig(d, e)
    assert np.max(np.abs(w - wr)) < 1e-12 * scale
    assert np.max(np.abs(T @ V - V * w)) < 1e-11 * scale
    assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-12


@pytest.mark.parametrize("n", [1, 3, 33, 129])
def test_random_tridiagonal(n):
    rng = np.random.default_rng(n)
    check(rng.uniform(-0.5, 1.5, n), rng.standard_normal(n - 1) * 0.3)


def test_clustered_triplets():
    """Tridiagonal form of Q diag(l) Q^T with triplets 1e-9 apart."""
    from scipy.linalg import hessenberg
    rng = np.random.default_rng(7)
    n = 96
    base = np.sort(rng.uniform(-1.0, 1.0, n // 3))
    lam = np.sort(np.concatenate([base, base + 1e-9, base + 2e-9]))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    T = hessenberg((q * lam) @ q.T)
    check(np.diag(T).copy(), np.diag(T, -1).copy())


def test_exactly_repeated_blocks():
    """20 uncoupled copies of one 3x3 block: every eigenvalue 20-fold."""
    d = np.tile([1.0, 2.0, 3.0], 20)
    e = np.tile([0.5, 0.5, 0.0], 20)[:-1]
    check(d, e)
