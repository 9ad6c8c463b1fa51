"""Parity at the Si216 workload: diamond Si216, 72^3 grid, 480 bands, 864
s+p projectors (the FP64 and mixed-precision sweep configurations).

At this size the FP32 stencil runs its 24-row tiles and the FP64 stencil its
12-row tiles over the full 72-plane z march, the grid-order nonlocal fill
sees every projector group of the cell, and the subspace eigensolve runs at
n = 480 (cooperative tridiagonal reduction on 120 workgroups, bisection and
inverse iteration, after the refinement attempt).  Reference unit: one
sweep of calc_sweep, sweep_module.f90:147-158.

Tolerances: hpsi 1e-12 (FP64) / 2e-6 (FP32) relative; eigenvalues of one
FP64 sweep 1e-10 Ha with the Ritz spans equal to 1e-9; after 8 mixed + 2
FP64 sweeps (FP32 or three-piece BF16 GEMMs) the occupied eigenvalues within
1e-6 Ha of 10 FP64 sweeps.
"""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32, PREC_MIXED_BF16

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def si216():
    from rsdft_amd.system import make_config
    return make_config("si216")


@pytest.fixture(scope="module")
def oracle216(si216):
    from oracle.oracle import Oracle
    return Oracle(si216)


@pytest.fixture(scope="module")
def start(si216, oracle216):
    """The seeded uniform(-1,1) start, orthonormalised by the oracle."""
    return oracle216.gram_schmidt(si216.psi0())


@pytest.fixture(scope="module")
def oracle_sweep(oracle216, start):
    print("oracle: one Si216 sweep (about 15 s of CPU)", flush=True)
    return oracle216.sweep(start)


@pytest.fixture(scope="module")
def dev(si216, gpu_ctx):
    d = RsdftGpu(si216, precision=PREC_FP64)
    yield d
    d.finalize()


@pytest.mark.parametrize("prec,tol", [(PREC_FP64, 1e-12), (PREC_MIXED_FP32, 2e-6)])
def test_hamiltonian_si216(dev, oracle216, si216, prec, tol):
    print(f"case: hamiltonian si216 16 bands precision {prec}", flush=True)
    psi = si216.psi0(100, 16)
    dev.set_precision(prec)
    try:
        got = dev.hamiltonian(psi)
    finally:
        dev.set_precision(PREC_FP64)
    err = rel(got, oracle216.hamiltonian(psi))
    print(f"  rel err {err:.3e}", flush=True)
    assert err < tol


@pytest.mark.parametrize("sweep_mode", [1, 0])
def test_one_sweep_si216_fp64(dev, oracle_sweep, start, si216, sweep_mode):
    """480-band CG + Gram-Schmidt + subspace diag, FP64, against the oracle."""
    print(f"case: one FP64 sweep si216 sweep mode {sweep_mode}", flush=True)
    xo, eo, ro = oracle_sweep
    dev.set_sweep_mode(sweep_mode)
    try:
        x, e, r = dev.sweep(start, 1)
    finally:
        dev.set_sweep_mode(1)
    de = float(np.max(np.abs(e - eo)))
    m = si216.grid.dV * (x @ xo.T)
    orth = float(np.max(np.abs(m.T @ m - np.eye(m.shape[0]))))
    print(f"  max |d eig| {de:.3e} Ha, span {orth:.3e}", flush=True)
    assert de < 1e-10
    assert orth < 1e-9
    assert np.all(np.diff(e) >= 0.0)


def _ritz_checks(x, e, si216, oracle216, orth_tol, ritz_tol):
    """Orthonormal x, and e are the Rayleigh quotients of x with the
    oracle's hamiltonian, which x diagonalises."""
    dV = si216.grid.dV
    S = dV * (x @ x.T)
    orth = float(np.max(np.abs(S - np.eye(S.shape[0]))))
    hx = oracle216.hamiltonian(x)
    Hs = dV * (x @ hx.T)
    scale = float(np.max(np.abs(e)))
    rq = float(np.max(np.abs(np.diag(Hs) - e)))
    off = float(np.max(np.abs(Hs - np.diag(np.diag(Hs)))))
    print(f"  orthonormality {orth:.2e}, Rayleigh quotients {rq:.2e} Ha, "
          f"off-diagonal {off:.2e} Ha", flush=True)
    assert orth < orth_tol
    assert rq < ritz_tol * scale
    assert off < ritz_tol * scale


@pytest.fixture(scope="module")
def fp64_ten(dev, start):
    """Ten FP64 sweeps from the start: the reference trajectory of the
    mixed-precision protocol."""
    y64 = start.copy()
    dev.set_precision(PREC_FP64)
    dev.attach(y64)
    _, e64, _ = dev.sweep(y64, 10)
    dev.sync()
    dev.detach()
    return y64, e64


@pytest.mark.parametrize("mode,name", [(PREC_MIXED_FP32, "fp32"), (PREC_MIXED_BF16, "bf16x3")])
def test_mixed_protocol_si216(dev, start, si216, oracle216, fp64_ten, mode, name):
    """8 mixed sweeps + 2 FP64 refinement sweeps against 10 FP64 sweeps from the
    same start (the FP64 trajectory's first sweep is the one checked against
    the oracle above), with FP32 or three-piece BF16 GEMMs.  The occupied
    eigenvalues are held to 1e-6 Ha.  The topmost unoccupied bands are not
    converged after ten sweeps: their values move by up to ~1e-5 Ha under any
    change of rounding along the trajectory (the FP64 run from the start
    rounded once to FP32 differs by 1.4e-6 Ha; summation-order variants of
    the mixed path gave 2e-6 to 1.1e-5 Ha), so all bands and the band energy
    get sanity bounds only."""
    print(f"case: si216 10 FP64 sweeps vs 8 mixed ({name}) + 2 FP64", flush=True)
    y64, e64 = fp64_ten
    ym = start.copy()
    dev.set_precision(mode)
    dev.attach(ym)
    dev.sweep(ym, 8)
    dev.sync()
    dev.detach()
    dev.set_precision(PREC_FP64)
    dev.attach(ym)
    _, em, _ = dev.sweep(ym, 2)
    dev.sync()
    dev.detach()
    d = np.abs(em - e64)
    occ = si216.occ > 0
    eband = abs(float(np.sum(si216.occ * (em - e64))))
    print(f"  max |d eig| {np.max(d):.3e} Ha (occupied {np.max(d[occ]):.3e}), "
          f"band energy {eband:.3e} Ha", flush=True)
    assert np.max(d[occ]) < 1e-6
    assert eband < 1e-5
    assert np.max(d) < 5e-5
    for y, e in ((y64, e64), (ym, em)):
        _ritz_checks(y, e, si216, oracle216, 1e-11, 1e-11)
