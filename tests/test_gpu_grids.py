"""Grids other than the cubic configurations, through the C-ABI against the
oracle: an anisotropic 20 x 24 x 28 cell (different stencil tiling in every
direction, ranges of the grid-order nonlocal fill that end inside planes)
with three atoms, for the hamiltonian in both precisions, the FP64 subspace
step, two FP64 sweeps, the mixed-precision sweeps, the density and the
nonlocal force; and a grid the stencil kernels cannot tile (x extent not a
multiple of 4), which must be refused at initialisation with an error.
Bounds as in the Si8 tests: H psi 1e-12 (FP64) and 2e-6 (FP32 storage)
relative, Ritz values 1e-10 Ha (1e-6 Ha after mixed sweeps), spans orthogonal
to 1e-9.
"""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, RsdftGpuError, PREC_FP64, PREC_MIXED_FP32
from rsdft_amd.system import make_custom

pytestmark = pytest.mark.gpu

ATOMS = np.array([[1.3, 2.1, 3.3], [4.6, 5.0, 7.9], [7.1, 8.2, 10.4]])


@pytest.fixture(scope="module")
def aniso():
    return make_custom((20, 24, 28), (8.55, 10.26, 11.97), ATOMS, nband=8)


@pytest.fixture(scope="module")
def aniso_oracle(aniso):
    from oracle.oracle import Oracle
    return Oracle(aniso)


@pytest.mark.parametrize("prec,name,tol", [(PREC_FP64, "FP64", 1e-12),
                                           (PREC_MIXED_FP32, "FP32", 2e-6)])
def test_hamiltonian_anisotropic(gpu_ctx, aniso, aniso_oracle, prec, name, tol):
    print(f"case: {name} hamiltonian, 20x24x28 cell, 3 atoms", flush=True)
    dev = RsdftGpu(aniso, precision=prec)
    psi = aniso.psi0(0, 8)
    h = dev.hamiltonian(psi)
    ref = aniso_oracle.hamiltonian(psi)
    err = float(np.max(np.abs(h - ref)) / np.max(np.abs(ref)))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < tol, err


def test_subspace_diag_anisotropic(gpu_ctx, aniso, aniso_oracle):
    print("case: FP64 subspace_diag, 20x24x28 cell, 3 atoms", flush=True)
    dev = RsdftGpu(aniso, precision=PREC_FP64)
    q = aniso_oracle.gram_schmidt(aniso.psi0())
    x, e = dev.subspace_diag(q)
    xo, eo, _ = aniso_oracle.subspace_diag(q)
    print(f"  max |d eig| {np.max(np.abs(e - eo)):.2e} Ha", flush=True)
    assert np.max(np.abs(e - eo)) < 1e-10


@pytest.mark.parametrize("pc_mode", [1, 0])
def test_two_sweeps_anisotropic(gpu_ctx, aniso, aniso_oracle, pc_mode):
    """CG (both preconditioner forms), Gram-Schmidt and subspace step on the
    anisotropic cell: Ritz values and the orthonormal span after each sweep."""
    print(f"case: 2 FP64 sweeps, 20x24x28 cell, precond mode {pc_mode}", flush=True)
    dev = RsdftGpu(aniso, precision=PREC_FP64)
    dev.set_precond_mode(pc_mode)
    x = aniso_oracle.gram_schmidt(aniso.psi0())
    xg = x.copy()
    for it in range(2):
        x, eo, _ = aniso_oracle.sweep(x)
        xg, eg, _ = dev.sweep(xg, 1)
        print(f"  sweep {it}: max|de| = {np.max(np.abs(eg - eo)):.3e}", flush=True)
        assert np.max(np.abs(eg - eo)) < 1e-10
    m = aniso.grid.dV * (xg @ x.T)
    assert np.max(np.abs(m.T @ m - np.eye(m.shape[0]))) < 1e-9


def test_mixed_sweeps_anisotropic(gpu_ctx, aniso, aniso_oracle):
    """4 FP32 sweeps + 2 FP64 sweeps against 6 FP64 oracle sweeps (< 1e-6 Ha)."""
    print("case: 4 mixed + 2 FP64 sweeps, 20x24x28 cell", flush=True)
    q = aniso_oracle.gram_schmidt(aniso.psi0())
    x = q
    for _ in range(6):
        x, eo, _ = aniso_oracle.sweep(x)
    dev = RsdftGpu(aniso, precision=PREC_MIXED_FP32)
    xm, em, _ = dev.sweep(q, 4)
    dev.set_precision(PREC_FP64)
    xm, em, _ = dev.sweep(xm, 2)
    err = float(np.max(np.abs(em - eo)))
    print(f"  max |eig err| = {err:.3e} Ha", flush=True)
    assert err < 1e-6, err


def test_density_and_force_anisotropic(gpu_ctx, aniso, aniso_oracle):
    """The adjoint outputs: density (1e-13) and nonlocal force (1e-11)."""
    print("case: FP64 density and nonlocal force, 20x24x28 cell", flush=True)
    dev = RsdftGpu(aniso, precision=PREC_FP64)
    q = aniso_oracle.gram_schmidt(aniso.psi0())
    rho = dev.calc_density(q, aniso.occ)
    ro = aniso_oracle.density(q, aniso.occ)
    f = dev.calc_force_ps_nloc2(q, aniso.occ)
    fo = aniso_oracle.force_nloc(q, aniso.occ)
    er = float(np.max(np.abs(rho - ro)) / np.max(np.abs(ro)))
    ef = float(np.max(np.abs(f - fo)) / np.max(np.abs(fo)))
    print(f"  density {er:.2e}, force {ef:.2e}", flush=True)
    assert er < 1e-13 and ef < 1e-11, (er, ef)


def test_untileable_grid_is_refused(gpu_ctx):
    """An x extent that is not a multiple of 4 cannot be tiled by the stencil
    kernels: initialisation reports it before anything is allocated."""
    print("case: 21x24x24 grid refused at init", flush=True)
    s = make_custom((21, 24, 24), (8.98, 10.26, 10.26), ATOMS[:1], nband=4)
    with pytest.raises(RsdftGpuError, match="multiple of 4"):
        RsdftGpu(s, precision=PREC_FP64)
