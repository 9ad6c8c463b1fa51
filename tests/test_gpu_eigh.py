"""The subspace eigensolvers (rsdft_gpu_test_eigh) against LAPACK's dsyevd
(numpy.linalg.eigh) on symmetric matrices shaped like Hsub.

Modes 2 and 3 are the LDS-resident Householder tridiagonalisation (one
cooperative launch) + the wave-per-columns back-transformation
(rocsolver_dormtr above n = 512), with the tridiagonal eigenpairs from
rocsolver_dstedc (2) or from bisection + inverse iteration (3; dstedc above
n = 512); mode 0 is rocsolver_dsyevd.  Bounds, with |H| the largest
|eigenvalue|: eigenvalues 1e-12 |H|, residual |H V - V W| 1e-11 |H|,
orthogonality |V^T V - I| 1e-12.
"""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(si8, gpu_ctx):
    return RsdftGpu(si8, precision=PREC_FP64)


def hsub_like(n, seed, mix=1e-2):
    """Ritz values spread over ~2 Ha plus a dense symmetric coupling."""
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(-0.5, 1.5, n))
    b = rng.standard_normal((n, n)) * mix
    return np.diag(d) + 0.5 * (b + b.T)


def clustered(n, seed):
    """Q diag(l) Q^T with triplets of eigenvalues 1e-9 apart."""
    rng = np.random.default_rng(seed)
    base = np.sort(rng.uniform(-1.0, 1.0, (n + 2) // 3))
    lam = np.sort(np.concatenate([base, base + 1e-9, base + 2e-9])[:n])
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (q * lam) @ q.T


def check(H, w, V):
    n = H.shape[0]
    wr = np.linalg.eigvalsh(H)
    scale = np.max(np.abs(wr))
    assert np.max(np.abs(w - wr)) < 1e-12 * scale
    assert np.max(np.abs(H @ V - V * w)) < 1e-11 * scale
    assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-12
    assert np.all(np.diag(V) >= 0.0)


@pytest.mark.parametrize("mode", [2, 3])
@pytest.mark.parametrize("n", [3, 5, 33, 64, 65, 129, 257, 480, 600])
def test_tridiagonal_eigensolver_shapes(dev, n, mode):
    print(f"case: test_eigh mode {mode} hsub_like n={n}", flush=True)
    H = hsub_like(n, seed=n)
    w, V, ms = dev.test_eigh(H, mode=mode)
    print(f"  {ms:.3f} ms", flush=True)
    check(H, w, V)


@pytest.mark.parametrize("mode", [3, 2, 0])
def test_eigensolver_clustered(dev, mode):
    print(f"case: test_eigh mode {mode} clustered n=480", flush=True)
    H = clustered(480, seed=7)
    w, V, ms = dev.test_eigh(H, mode=mode)
    print(f"  {ms:.3f} ms", flush=True)
    check(H, w, V)


@pytest.mark.parametrize("mode", [2, 3])
def test_tridiagonal_matches_dsyevd(dev, mode):
    """Same Ritz pairs as rocsolver_dsyevd (vectors up to the shared sign
    convention), at the Si216 subspace size."""
    print(f"case: test_eigh mode {mode} vs mode 0 n=480", flush=True)
    H = hsub_like(480, seed=11, mix=3e-2)
    w2, V2, t2 = dev.test_eigh(H, mode=mode)
    w0, V0, t0 = dev.test_eigh(H, mode=0)
    print(f"  mode {mode} {t2:.3f} ms, dsyevd {t0:.3f} ms", flush=True)
    assert np.max(np.abs(w2 - w0)) < 1e-12
    gap = np.min(np.diff(w0))
    assert gap > 1e-6
    assert np.max(np.abs(V2 - V0)) < 1e-8


def test_tridiagonal_solver_exact_degeneracy(dev):
    """Exactly repeated eigenvalues (20 copies of one 3x3 block, uncoupled):
    bisection + inverse iteration must return an orthonormal basis of each
    degenerate eigenspace."""
    print("case: test_eigh mode 3 degenerate blocks n=60", flush=True)
    blk = np.array([[1.0, 0.5, 0.0], [0.5, 2.0, 0.5], [0.0, 0.5, 3.0]])
    H = np.kron(np.eye(20), blk)
    w, V, ms = dev.test_eigh(H, mode=3)
    check(H, w, V)


def test_tridiagonal_solver_mixed_clusters(dev):
    """Half the spectrum well separated, half in exact pairs: workgroups with
    and without cluster members in one solve (the former finish all
    inverse-iteration rounds in one launch, the latter orthogonalise between
    rounds)."""
    print("case: test_eigh mode 3 mixed separated / paired spectrum n=300", flush=True)
    rng = np.random.default_rng(5)
    single = rng.uniform(-1.0, 0.0, 150)
    pairs = rng.uniform(0.0, 1.0, 75)
    lam = np.sort(np.concatenate([single, pairs, pairs]))
    q, _ = np.linalg.qr(rng.standard_normal((300, 300)))
    H = (q * lam) @ q.T
    H = 0.5 * (H + H.T)
    w, V, ms = dev.test_eigh(H, mode=3)
    check(H, w, V)


def nearly_diagonal(n, seed, coupling=1e-4, ncluster=20, csize=3, csplit=1e-6):
    """Ritz values spread over 2 Ha with clusters of csize values csplit apart,
    plus a symmetric coupling of size `coupling`: the Hsub of bands that are
    close to converged, the case the refinement eigensolver (mode 1) is for."""
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(-0.5, 1.5, n))
    cand = np.arange(0, n - csize, 2 * csize)
    starts = rng.choice(cand, min(ncluster, len(cand) // 2), replace=False)
    for s in starts:
        d[s:s + csize] = d[s] + csplit * np.arange(csize)
    b = rng.standard_normal((n, n)) * coupling
    return np.diag(d) + 0.5 * (b + b.T)


@pytest.mark.parametrize("n", [480, 96])
def test_refinement_nearly_diagonal(dev, n):
    """Mode 1 on a nearly diagonal Hsub with near-degenerate clusters: the
    refinement path solves it (counted in the timers) and matches dsyevd."""
    print(f"case: test_eigh mode 1 nearly diagonal n={n}", flush=True)
    H = nearly_diagonal(n, seed=n)
    dev.reset_timers()
    w, V, ms = dev.test_eigh(H, mode=1)
    used = dev.timers()["eig_refined"]
    print(f"  {ms:.3f} ms, refined {used:.0f}", flush=True)
    assert used == 1
    check(H, w, V)


def test_refinement_pair_overflow_falls_back(dev):
    """More strong pairs than the cluster search holds (a block of 72 coupled
    near-degenerate values: 2556 pairs > 2048): mode 1 must fall back to the
    tridiagonal solver and still return the right eigenpairs."""
    print("case: test_eigh mode 1 strong-pair overflow n=480", flush=True)
    n = 480
    H = nearly_diagonal(n, seed=3, ncluster=0)
    H[100:172, 100:172] += 1e-3 * np.random.default_rng(4).standard_normal((72, 72))
    H[100:172, 100:172] = 0.5 * (H[100:172, 100:172] + H[100:172, 100:172].T)
    np.fill_diagonal(H[100:172, 100:172], H[100, 100])
    dev.reset_timers()
    w, V, ms = dev.test_eigh(H, mode=1)
    used = dev.timers()["eig_refined"]
    print(f"  {ms:.3f} ms, refined {used:.0f}", flush=True)
    assert used == 0
    check(H, w, V)


@pytest.mark.parametrize("mode", [1, 0, 3])
def test_eigensolver_n1_and_n2(dev, mode):
    """The smallest subspaces: n = 1 and n = 2."""
    print(f"case: test_eigh mode {mode} n=1,2", flush=True)
    for H in (np.array([[0.7]]), np.array([[0.3, 0.1], [0.1, -0.2]])):
        w, V, ms = dev.test_eigh(H, mode=mode)
        check(H, w, V)
