"""The GPU CG through the band-choice and reject rules (cg_module.f90:319-323,
361-364), on the plane-wave cases of tests/planewave.py where the oracle is
pinned to a model of the step: band 0 swaps to the upper Ritz pair, band 1
takes its step, band 2's step is rejected (its x and E are restored).  The
rejected band's update is undone on the GPU (k_cg_accept) and, in a fused
sweep, its H x is recomputed before the Rayleigh-Ritz step (run_sweeps)."""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32
from tests.planewave import PlaneWaveCell, pinning_cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from oracle.oracle import Oracle
    cell = PlaneWaveCell()
    X = np.stack([cell.vector(b, c) for b, c in pinning_cases()])
    o1 = Oracle(cell.system, Ncg=1, mloop=3)
    x1, e1, _, cnt1 = o1.conjugate_gradient(X, counts=True)
    o3 = Oracle(cell.system)
    x3, e3, _, cnt3 = o3.conjugate_gradient(X, counts=True)
    assert cnt1[1] == 1 and cnt3[1] == 1
    return cell, X, (x1, e1), (x3, e3), o3


@pytest.fixture(scope="module")
def dev(case, gpu_ctx):
    cell = case[0]
    return RsdftGpu(cell.system, precision=PREC_FP64, with_derivatives=False)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("pc_mode", [1, 0])
@pytest.mark.parametrize("ncg", [1, 3])
def test_cg_rules_fp64(dev, case, pc_mode, ncg):
    print(f"case: CG rules fp64 ncg={ncg} precond mode {pc_mode}", flush=True)
    cell, X, r1, r3, _ = case
    xo, eo = r1 if ncg == 1 else r3
    dev.set_precision(PREC_FP64)
    dev.set_precond_mode(pc_mode)
    dev._call("rsdft_gpu_set_solver_params", ncg, 3)
    try:
        x, e, _ = dev.conjugate_gradient(X)
        rejected = dev.timers()["cg_rejected_bands"]
    finally:
        dev._call("rsdft_gpu_set_solver_params", 3, 3)
        dev.set_precond_mode(1)
    print(f"  eig {e} oracle {eo} rejected bands {rejected}", flush=True)
    assert rejected == 1
    assert np.max(np.abs(e - eo)) < 1e-10
    assert rel(x, xo) < 1e-7
    if ncg == 1:
        assert rel(x[2], X[2]) < 1e-12   # the undone update


def test_cg_rules_mixed(dev, case):
    print("case: CG rules mixed fp32 ncg=1", flush=True)
    cell, X, (xo, eo), _, _ = case
    dev.set_precision(PREC_MIXED_FP32)
    dev._call("rsdft_gpu_set_solver_params", 1, 3)
    try:
        x, e, _ = dev.conjugate_gradient(X)
        rejected = dev.timers()["cg_rejected_bands"]
    finally:
        dev._call("rsdft_gpu_set_solver_params", 3, 3)
        dev.set_precision(PREC_FP64)
    print(f"  eig {e} oracle {eo} rejected bands {rejected}", flush=True)
    assert rejected == 1
    assert np.max(np.abs(e - eo)) < 1e-6
    assert rel(x[2], X[2]) < 1e-6


@pytest.mark.parametrize("sweep_mode", [1, 0])
def test_sweep_with_rejected_band(dev, case, sweep_mode):
    """CG (Ncg = 3, one band rejected) + Gram-Schmidt + subspace diag."""
    print(f"case: sweep with a rejected CG band, sweep mode {sweep_mode}", flush=True)
    cell, X, _, _, o3 = case
    xo, eo, _ = o3.sweep(X)
    dev.set_precision(PREC_FP64)
    dev.set_sweep_mode(sweep_mode)
    try:
        x, e, _ = dev.sweep(X, 1)
        rejected = dev.timers()["cg_rejected_bands"]
    finally:
        dev.set_sweep_mode(1)
    print(f"  eig {e} oracle {eo} rejected bands {rejected}", flush=True)
    assert rejected == 1
    assert np.max(np.abs(e - eo)) < 1e-10
    dV = cell.grid.dV
    # the same Ritz vectors up to sign
    ov = np.abs(np.sum(x * xo, axis=1) * dV)
    assert np.max(np.abs(ov - 1.0)) < 1e-9
