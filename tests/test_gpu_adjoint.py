"""Adjoint (backward) entry through the C-ABI: dE/dVloc and dE/dR of the band
energy E = sum_n f_n <psi_n|H|psi_n> dV at fixed psi, on the Si64 adjoint
configuration (48^3, 160 bands, atom 1 displaced by (0.1, 0.05, 0) bohr).

Checked against central finite differences of the GPU's own band energy
(step 1e-4, relative 1e-6; the projector point sets are frozen at the
unmoved atom, the function whose derivative the nonlocal force is) and
against the oracle's density and nonlocal force at the same psi."""
import numpy as np
import pytest

from rsdft_amd import system as S
from rsdft_amd.gpu import RsdftGpu, PREC_FP64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def si64():
    return S.make_config("si64")


@pytest.fixture(scope="module")
def setup(si64, gpu_ctx):
    from oracle.oracle import Oracle
    dev = RsdftGpu(si64, precision=PREC_FP64)
    x = dev.gram_schmidt(si64.psi0())
    return dev, x, Oracle(si64)


def tables_at(s, atom, shift):
    delta = s.atom_delta.copy()
    delta[atom] += shift
    return S.make_projectors(s.grid, s.atom_ic, delta, s.D, sphere_delta=s.atom_delta)


def test_dE_dV_matches_density_and_fd(setup, si64):
    print("case: dE/dV si64 160 bands", flush=True)
    dev, x, o = setup
    E, dEdV, dEdR = dev.band_energy_grad(x, si64.occ)
    rho = o.density(x, si64.occ)
    assert np.max(np.abs(dEdV - si64.grid.dV * rho)) < 1e-13 * np.max(np.abs(dEdV)) + 1e-18
    assert abs(E - o.band_energy(x, si64.occ)) < 1e-10 * abs(E)
    rng = np.random.default_rng(11)
    for r in rng.choice(si64.grid.ML, 3, replace=False):
        v = si64.vloc.copy()
        v[r] += 1e-4
        dev.set_vloc(v)
        Ep = dev.band_energy_grad(x, si64.occ)[0]
        v[r] -= 2e-4
        dev.set_vloc(v)
        Em = dev.band_energy_grad(x, si64.occ)[0]
        dev.set_vloc(si64.vloc)
        fd = (Ep - Em) / 2e-4
        # E is linear in V: the difference is exact up to the rounding of E
        assert abs(fd - dEdV[r]) <= 1e-6 * abs(dEdV[r]) + 1e-15 * abs(E) / 1e-4, (r, fd, dEdV[r])


def nonlocal_energy(dev, x, occ, dV):
    """E_NL = sum_n f_n <psi_n|V_NL|psi_n> dV through rsdft_gpu_op_nonlocal; only this
    part of E depends on the atomic positions (Vloc is a fixed field here)."""
    vx = dev.op_nonlocal(x)
    return float(np.sum(occ * np.einsum("ij,ij->i", x, vx)) * dV)


def test_dE_dR_matches_force_oracle_and_fd(setup, si64):
    print("case: dE/dR si64 160 bands", flush=True)
    dev, x, o = setup
    _, _, dEdR = dev.band_energy_grad(x, si64.occ)
    F = o.force_nloc(x, si64.occ)
    assert np.max(np.abs(dEdR + F)) < 1e-10 * np.max(np.abs(F))
    occ, dV = si64.occ, si64.grid.dV
    for a in (0, 1, 37, 63):
        for d in range(3):
            e = np.zeros(3)
            e[d] = 1e-4
            dev.set_projectors(tables_at(si64, a, e), with_derivatives=False)
            Ep = nonlocal_energy(dev, x, occ, dV)
            dev.set_projectors(tables_at(si64, a, -e), with_derivatives=False)
            Em = nonlocal_energy(dev, x, occ, dV)
            fd = (Ep - Em) / 2e-4
            print(f"  atom {a} dir {d}: analytic {dEdR[a, d]:.10e} fd {fd:.10e}", flush=True)
            assert abs(fd - dEdR[a, d]) < 1e-6 * max(abs(dEdR[a, d]), 1e-3 * np.max(np.abs(dEdR))), (a, d)
    dev.set_projectors(si64.proj)
