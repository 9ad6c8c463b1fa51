"""Parity of the HIP library (through the C-ABI) with the oracle, FP64, Si8.

Tolerances: hpsi 1e-12 relative (max norm), eigenvalues 1e-10 Ha, as the
FP64 path must match the reference's CPU path.
"""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def dev(si8, gpu_ctx):
    return RsdftGpu(si8, precision=PREC_FP64)


@pytest.fixture(scope="module")
def psi8(si8):
    return si8.psi0(0, 8)


def test_hamiltonian_fp64(dev, oracle_si8, psi8):
    print("case: hamiltonian fp64 si8 8 bands", flush=True)
    assert rel(dev.hamiltonian(psi8), oracle_si8.hamiltonian(psi8)) < 1e-12


def test_op_kinetic_overwrites(dev, oracle_si8, psi8):
    print("case: op_kinetic fp64 si8", flush=True)
    assert rel(dev.op_kinetic(psi8), oracle_si8.op_kinetic(psi8)) < 1e-12


def test_op_localpot_accumulates(dev, oracle_si8, psi8, si8):
    print("case: op_localpot fp64 si8", flush=True)
    base = si8.psi0(8, 8)
    got = dev.op_localpot(psi8, base)
    want = oracle_si8.op_localpot(psi8, base.copy())
    assert rel(got, want) < 1e-13


def test_op_nonlocal_accumulates(dev, oracle_si8, psi8, si8):
    print("case: op_nonlocal fp64 si8", flush=True)
    base = si8.psi0(8, 8)
    got = dev.op_nonlocal(psi8, base)
    want = oracle_si8.op_nonlocal(psi8, base.copy())
    assert rel(got, want) < 1e-12


def test_gram_schmidt_fp64(dev, oracle_si8, si8):
    print("case: gram_schmidt fp64 si8 32 bands", flush=True)
    psi = si8.psi0()
    q = dev.gram_schmidt(psi)
    qo = oracle_si8.gram_schmidt(psi)
    assert np.max(np.abs(q @ q.T * si8.grid.dV - np.eye(psi.shape[0]))) < 1e-12
    assert rel(q, qo) < 1e-11


def test_gram_schmidt_fp64_many_blocks(dev, oracle_si8, si8):
    """More bands than one 120-column diagonal block of the blocked Cholesky
    (the block rows go through dtrsm and dsyrk): 130 bands on the Si8 grid."""
    print("case: gram_schmidt fp64 si8 130 bands", flush=True)
    psi = si8.psi0(0, 130)
    q = dev.gram_schmidt(psi)
    qo = oracle_si8.gram_schmidt(psi)
    assert np.max(np.abs(q @ q.T * si8.grid.dV - np.eye(psi.shape[0]))) < 1e-12
    assert rel(q, qo) < 1e-10


def test_gram_schmidt_zero_band_is_refused(dev, si8):
    """A zero band makes the overlap singular: its pivot of the Cholesky
    factorisation is exactly zero, which the library reports as an error (the
    reference's dpotrf returns info = j + 1 there) instead of returning
    non-orthonormal vectors."""
    from rsdft_amd.gpu import RsdftGpuError
    print("case: gram_schmidt fp64 si8 zero band", flush=True)
    psi = si8.psi0()
    psi[5] = 0.0
    with pytest.raises(RsdftGpuError, match="Cholesky"):
        dev.gram_schmidt(psi)
    q = dev.gram_schmidt(si8.psi0())   # the context is usable afterwards
    assert np.max(np.abs(q @ q.T * si8.grid.dV - np.eye(q.shape[0]))) < 1e-12


@pytest.mark.parametrize("eig_mode", [1, 0, 2, 3])
def test_subspace_diag_fp64(dev, oracle_si8, si8, eig_mode):
    """eig_mode 1 (refinement, here falling back to mode 3: Hsub of random
    vectors is far from diagonal), 0 rocsolver_dsyevd, 2 the LDS-resident
    tridiagonal reduction + dstedc, 3 the reduction + bisection / inverse
    iteration."""
    print(f"case: subspace_diag fp64 si8 eigensolver mode {eig_mode}", flush=True)
    q = oracle_si8.gram_schmidt(si8.psi0())
    dev.set_eigensolver(eig_mode)
    try:
        x, e = dev.subspace_diag(q)
    finally:
        dev.set_eigensolver(1)
    xo, eo, _ = oracle_si8.subspace_diag(q)
    assert np.max(np.abs(e - eo)) < 1e-10
    # Ritz vectors agree up to sign
    s = np.sign(np.sum(x * xo, axis=1))
    assert rel(x * s[:, None], xo) < 1e-8


@pytest.mark.parametrize("pc_mode", [1, 0])
def test_conjugate_gradient_fp64(dev, oracle_si8, si8, pc_mode):
    print(f"case: conjugate_gradient fp64 si8 precond mode {pc_mode}", flush=True)
    q = oracle_si8.gram_schmidt(si8.psi0())
    q, _, _ = oracle_si8.subspace_diag(q)
    dev.set_precond_mode(pc_mode)
    x, e, r = dev.conjugate_gradient(q)
    dev.set_precond_mode(1)
    xo, eo, ro = oracle_si8.conjugate_gradient(q)
    assert np.max(np.abs(e - eo)) < 1e-10
    assert np.max(np.abs(r - ro) / np.maximum(ro, 1e-30)) < 1e-8
    assert rel(x, xo) < 1e-9


@pytest.mark.parametrize("pc_mode,sweep_mode", [(1, 1), (0, 1), (1, 0)])
def test_four_sweeps_fp64(dev, oracle_si8, si8, pc_mode, sweep_mode):
    print(f"case: 4 sweeps fp64 si8 precond mode {pc_mode} sweep mode {sweep_mode}", flush=True)
    dev.set_precond_mode(pc_mode)
    dev.set_sweep_mode(sweep_mode)
    q = oracle_si8.gram_schmidt(si8.psi0())
    x, eo = q, None
    xg = q.copy()
    try:
        for it in range(4):
            x, eo, ro = oracle_si8.sweep(x)
            xg, eg, rg = dev.sweep(xg, 1)
            print(f"  sweep {it}: max|de| = {np.max(np.abs(eg - eo)):.3e}", flush=True)
            assert np.max(np.abs(eg - eo)) < 1e-10
        # the same orthonormal span: the overlap dV X_gpu X_oracle^T is orthogonal
        m = oracle_si8.dV * (xg @ x.T)
        assert np.max(np.abs(m.T @ m - np.eye(m.shape[0]))) < 1e-9
    finally:
        dev.set_precond_mode(1)
        dev.set_sweep_mode(1)


def test_density_and_force_fp64(dev, oracle_si8, si8):
    print("case: density/force fp64 si8", flush=True)
    q = oracle_si8.gram_schmidt(si8.psi0())
    rho = dev.calc_density(q, si8.occ)
    assert rel(rho, oracle_si8.density(q, si8.occ)) < 1e-13
    f = dev.calc_force_ps_nloc2(q, si8.occ)
    fo = oracle_si8.force_nloc(q, si8.occ)
    assert rel(f, fo) < 1e-11


def test_refinement_eigensolver_matches_dsyevd(dev, oracle_si8, si8):
    """Nearly diagonal Hsub (Ritz vectors after 6 sweeps, one CG + GS): the
    refinement path and rocsolver_dsyevd give the same Ritz pairs."""
    print("case: refinement eigensolver vs dsyevd si8", flush=True)
    x = oracle_si8.gram_schmidt(si8.psi0())
    for _ in range(6):
        x, _, _ = oracle_si8.sweep(x)
    x, _, _ = oracle_si8.conjugate_gradient(x)
    q = oracle_si8.gram_schmidt(x)
    xo, eo, _ = oracle_si8.subspace_diag(q)
    dev.reset_timers()
    dev.set_eigensolver(1)
    x1, e1 = dev.subspace_diag(q)
    used = dev.timers()["eig_refined"]
    dev.set_eigensolver(0)
    x0, e0 = dev.subspace_diag(q)
    dev.set_eigensolver(2)
    x2, e2 = dev.subspace_diag(q)
    dev.set_eigensolver(1)
    assert used == 1
    for e in (e1, e0, e2):
        assert np.max(np.abs(e - eo)) < 1e-10
    for x in (x1, x2):
        s = np.sign(np.sum(x * xo, axis=1))
        assert rel(x * s[:, None], xo) < 1e-8
