"""Mixed-precision path (FP32 storage and operators; FP32 or BF16x3 MFMA
GEMMs with FP64-flushed accumulation; FP64 Cholesky, eigensolve, CG
scalars) through the C-ABI, checked against the FP64 oracle on Si8."""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16

pytestmark = pytest.mark.gpu

MODES = [(PREC_MIXED_FP32, "fp32"), (PREC_MIXED_BF16, "bf16x3")]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def dev(si8, gpu_ctx):
    return RsdftGpu(si8, precision=PREC_FP64)


@pytest.mark.parametrize("mode,name", MODES)
def test_hamiltonian_mixed(dev, oracle_si8, si8, mode, name):
    print(f"case: hamiltonian {name} si8", flush=True)
    dev.set_precision(mode)
    psi = si8.psi0(0, 8)
    err = rel(dev.hamiltonian(psi), oracle_si8.hamiltonian(psi))
    dev.set_precision(PREC_FP64)
    assert err < 2e-6, err


@pytest.mark.parametrize("mode,name", MODES)
def test_gram_schmidt_mixed(dev, si8, mode, name):
    print(f"case: gram_schmidt {name} si8", flush=True)
    dev.set_precision(mode)
    q = dev.gram_schmidt(si8.psi0())
    dev.set_precision(PREC_FP64)
    orth = np.max(np.abs(q @ q.T * si8.grid.dV - np.eye(q.shape[0])))
    assert orth < 1e-5, orth


@pytest.mark.parametrize("mode,name", MODES)
def test_subspace_diag_mixed(dev, oracle_si8, si8, mode, name):
    print(f"case: subspace_diag {name} si8", flush=True)
    q = oracle_si8.gram_schmidt(si8.psi0())
    _, eo, _ = oracle_si8.subspace_diag(q)
    dev.set_precision(mode)
    _, e = dev.subspace_diag(q)
    dev.set_precision(PREC_FP64)
    assert np.max(np.abs(e - eo)) < 5e-5


@pytest.mark.parametrize("mode,name", MODES)
def test_mixed_sweeps_then_fp64_refinement(dev, oracle_si8, si8, mode, name):
    """4 mixed sweeps + 2 FP64 sweeps against 6 FP64 oracle sweeps."""
    print(f"case: mixed sweeps {name} si8", flush=True)
    q = oracle_si8.gram_schmidt(si8.psi0())
    x = q
    for _ in range(6):
        x, eo, _ = oracle_si8.sweep(x)
    dev.set_precision(mode)
    xm, em, _ = dev.sweep(q, 4)
    dev.set_precision(PREC_FP64)
    xm, em, _ = dev.sweep(xm, 2)
    err = float(np.max(np.abs(em - eo)))
    print(f"  max |eig err| = {err:.3e} Ha", flush=True)
    assert err < 1e-6, err


def test_bf16x3_products_match_fp32_products(si8, gpu_ctx):
    """The three-piece BF16 split (six MFMA products per block) is as
    accurate as the FP32 matrix-core products: on the Si64 grid (110,592-point
    reductions, 160 bands) the orthogonality defect of Gram-Schmidt (S = dV
    psi^T psi, then psi R^-1) and the subspace Ritz values against the FP64
    path stay within a small factor of what the FP32 products give."""
    from rsdft_amd.system import make_config
    s = make_config("si64")
    x0 = s.psi0()
    dev = RsdftGpu(s, precision=PREC_FP64)
    try:
        q64 = dev.gram_schmidt(x0)
        _, e64 = dev.subspace_diag(q64)
        orth, eerr = {}, {}
        for mode, name in MODES:
            print(f"case: split products {name} si64", flush=True)
            dev.set_precision(mode)
            q = dev.gram_schmidt(x0)
            orth[name] = float(np.max(np.abs(q @ q.T * s.grid.dV - np.eye(q.shape[0]))))
            _, e = dev.subspace_diag(q64)
            eerr[name] = float(np.max(np.abs(e - e64)))
            dev.set_precision(PREC_FP64)
            print(f"  orthogonality {orth[name]:.3e}, ritz values {eerr[name]:.3e} Ha", flush=True)
        assert orth["fp32"] < 1e-5
        assert orth["bf16x3"] < max(4 * orth["fp32"], 1e-7)
        assert eerr["bf16x3"] < max(4 * eerr["fp32"], 1e-7)
    finally:
        dev.finalize()
