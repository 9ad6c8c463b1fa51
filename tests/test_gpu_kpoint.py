"""Complex wavefunctions at one k point on the GPU (csrc/kpoint.hip) against
the oracle's complex restatement (oracle/kpoint.py, pinned by
tests/test_oracle_kpoint.py), on a two-atom cell whose projector spheres
cross the cell boundary (the Bloch phases of folded images matter) and on
Si8:

* the hamiltonian with zcoef_kin and Bloch-phased projectors: 1e-12 relative;
  at k = 0 on real wavefunctions it equals the real hamiltonian;
* conjugate_gradient (zhegv 2x2 steps, phase fix, preconditioner with
  Hermitian inner products), Gram-Schmidt (orthonormal, positive real R
  diagonal: the same Q), subspace_diag (zheevd) and three sweeps:
  eigenvalues 1e-10 Ha (sweeps 1e-9), the same spans;
* time reversal: the sweep at -k from conj(psi) gives the eigenvalues at k.
"""
import numpy as np
import pytest

from oracle.kpoint import KpointOracle
from rsdft_amd.gpu import RsdftGpu, PREC_FP64
from rsdft_amd.system import kpoint_kinetic, make_custom

pytestmark = pytest.mark.gpu

K = np.array([0.11, -0.07, 0.05])


def _oracle(system, kvec):
    ck2, nabk = kpoint_kinetic(system.Md, system.grid.h, kvec)
    return KpointOracle(system, system.kpoint_projectors(kvec), ck2, nabk)


@pytest.fixture(scope="module", params=["cell2", "si8"])
def case(request, gpu_ctx):
    if request.param == "cell2":
        pos = np.array([[0.3, 0.2, 0.1], [4.1, 3.0, 2.6]])
        s = make_custom((16, 12, 12), (6.4, 5.2, 5.6), pos, Md=4, nband=6, seed=11)
    else:
        from rsdft_amd.system import make_config
        s = make_config("si8", nband=8)
    dev = RsdftGpu(s, precision=PREC_FP64)
    dev.set_kpoint(K, s.kpoint_projectors(K))
    yield request.param, s, dev
    dev.finalize()


def _span_gap(a, b, dV):
    """Largest deviation of |<a|b>| dV from a unitary matrix (same span)."""
    S = np.conj(a) @ b.T * dV
    return float(np.max(np.abs(S @ np.conj(S.T) - np.eye(len(S)))))


def test_zhamiltonian_matches_oracle(case):
    name, s, dev = case
    print(f"case: zhamiltonian {name} k={K}", flush=True)
    zo = _oracle(s, K)
    x = s.psi0c()
    h = dev.zhamiltonian(x)
    ho = zo.hamiltonian(x)
    err = np.max(np.abs(h - ho)) / np.max(np.abs(ho))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < 1e-12


def test_zhamiltonian_k0_is_the_real_hamiltonian(case):
    name, s, dev = case
    print(f"case: zhamiltonian {name} k=0 vs real", flush=True)
    # (one library context: the k point is switched on the same one)
    dev.set_kpoint(np.zeros(3), s.kpoint_projectors(np.zeros(3)))
    try:
        x = s.psi0()
        hz = dev.zhamiltonian(x)
        hr = dev.hamiltonian(x)
        assert np.max(np.abs(hz.imag)) == 0.0
        assert np.max(np.abs(hz.real - hr)) < 1e-13 * np.max(np.abs(hr))
    finally:
        dev.set_kpoint(K, s.kpoint_projectors(K))


def test_zeigensolver_pieces_match_oracle(case):
    name, s, dev = case
    print(f"case: zcg / zgs / zsd {name}", flush=True)
    zo = _oracle(s, K)
    dV = s.grid.dV
    x0 = zo.gram_schmidt(s.psi0c())
    xg, eg, rg = dev.zconjugate_gradient(x0)
    xo, eo, ro = zo.conjugate_gradient(x0)
    print(f"  cg eig {np.max(np.abs(eg - eo)):.2e}, x {np.max(np.abs(xg - xo)):.2e}", flush=True)
    assert np.max(np.abs(eg - eo)) < 1e-10
    assert np.max(np.abs(xg - xo)) < 1e-8 * np.max(np.abs(xo))
    assert np.max(np.abs(rg - ro)) < 1e-8 * max(1.0, np.max(np.abs(ro)))
    qg = dev.zgram_schmidt(xo)
    qo = zo.gram_schmidt(xo)
    print(f"  gs {np.max(np.abs(qg - qo)):.2e}", flush=True)
    assert np.max(np.abs(qg - qo)) < 1e-10 * np.max(np.abs(qo))
    S = np.conj(qg) @ qg.T * dV
    assert np.max(np.abs(S - np.eye(len(S)))) < 1e-12
    vg, wg = dev.zsubspace_diag(qo)
    vo, wo, _ = zo.subspace_diag(qo)
    print(f"  sd eig {np.max(np.abs(wg - wo)):.2e}", flush=True)
    assert np.max(np.abs(wg - wo)) < 1e-11
    assert _span_gap(vg, vo, dV) < 1e-9


def test_zsweeps_match_oracle_and_time_reversal(case):
    name, s, dev = case
    print(f"case: 3 zsweeps {name}", flush=True)
    zo = _oracle(s, K)
    x0 = zo.gram_schmidt(s.psi0c())
    xg, eg, _ = dev.zsweep(x0, 3)
    xo = x0
    for _ in range(3):
        xo, eo, _ = zo.sweep(xo)
    print(f"  eig {np.max(np.abs(eg - eo)):.2e}", flush=True)
    assert np.max(np.abs(eg - eo)) < 1e-9
    assert np.all(np.diff(eg) >= -1e-12)
    dev.set_kpoint(-K, s.kpoint_projectors(-K))
    try:
        _, em, _ = dev.zsweep(np.conj(x0), 3)
    finally:
        dev.set_kpoint(K, s.kpoint_projectors(K))
    print(f"  -k vs k {np.max(np.abs(em - eg)):.2e}", flush=True)
    assert np.max(np.abs(em - eg)) < 1e-10
