"""Gradient-check harness: the adjoint entry rsdft_gpu_band_energy_grad wrapped
in a torch.autograd.Function (test-only; the product is the C-ABI entry).

E(V_loc, R) = sum_n f_n <psi_n|H|psi_n> dV at fixed psi on the Si64 adjoint
configuration (48^3, 160 bands, atom 1 displaced by (0.1, 0.05, 0) bohr).
Forward sets V_loc and the projector tables of the positions R (point sets
frozen at the unmoved atoms, the function whose R-derivative the nonlocal
force is) and returns E; backward returns dE/dV = dV rho
(density_module.f90:129-136) and dE/dR = -F_nl (ps_nloc2_module.f90:1601-1773).

Checked with torch.autograd.gradcheck along random directions in (V, R)
(central differences of the forward, step 1e-4, relative 1e-6), and the
gradients against the oracle's density and nonlocal force at the same psi.
"""
import numpy as np
import pytest

from rsdft_amd import system as S
from rsdft_amd.gpu import RsdftGpu, PREC_FP64

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class BandEnergy(torch.autograd.Function):
    """E(V_loc, R) through the GPU library (CPU float64 tensors in and out)."""

    @staticmethod
    def forward(ctx, vloc, pos, dev, system, x):
        h = np.array(system.grid.h)
        delta = pos.detach().numpy() - system.atom_ic * h[None, :]
        dev.set_vloc(vloc.detach().numpy())
        dev.set_projectors(S.make_projectors(system.grid, system.atom_ic, delta, system.D,
                                             sphere_delta=system.atom_delta))
        E, dEdV, dEdR = dev.band_energy_grad(x, system.occ)
        ctx.save_for_backward(torch.from_numpy(dEdV), torch.from_numpy(dEdR))
        return torch.tensor(E, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        dEdV, dEdR = ctx.saved_tensors
        return gE * dEdV, gE * dEdR, None, None, None


@pytest.fixture(scope="module")
def setup(gpu_ctx):
    from oracle.oracle import Oracle
    si64 = S.make_config("si64")
    dev = RsdftGpu(si64, precision=PREC_FP64)
    x = dev.gram_schmidt(si64.psi0())
    return si64, dev, x, Oracle(si64)


def test_gradients_match_density_and_force(setup):
    print("case: autograd harness gradients vs oracle density/force si64", flush=True)
    s, dev, x, o = setup
    v = torch.tensor(s.vloc, requires_grad=True)
    R = torch.tensor(s.positions, requires_grad=True)
    E = BandEnergy.apply(v, R, dev, s, x)
    gv, gR = torch.autograd.grad(E, (v, R))
    rho = o.density(x, s.occ)
    F = o.force_nloc(x, s.occ)
    assert abs(E.item() - o.band_energy(x, s.occ)) < 1e-10 * abs(E.item())
    assert np.max(np.abs(gv.numpy() - s.grid.dV * rho)) < 1e-13 * np.max(np.abs(gv.numpy()))
    assert np.max(np.abs(gR.numpy() + F)) < 1e-10 * np.max(np.abs(F))


@pytest.mark.parametrize("seed", [1, 2])
def test_gradcheck_along_random_directions(setup, seed):
    """gradcheck of t -> E(V + t dV, R + t dR) at t = 0: its backward (the
    adjoint kernels) against central differences of its forward."""
    print(f"case: autograd gradcheck si64 direction seed {seed}", flush=True)
    s, dev, x, o = setup
    rng = np.random.default_rng(seed)
    dv = torch.tensor(rng.standard_normal(s.grid.ML) * 0.05)
    dR = torch.tensor(rng.standard_normal((s.natom, 3)) * 0.05)
    v0 = torch.tensor(s.vloc)
    R0 = torch.tensor(s.positions)

    def f(t):
        return BandEnergy.apply(v0 + t * dv, R0 + t * dR, dev, s, x)

    t = torch.zeros(1, dtype=torch.float64, requires_grad=True)
    # central differences of E ~ 6.3e3 Ha with step 1e-4: the rounding of the
    # two FP64 band-energy sums alone allows ~ 50 eps |E| / 1e-4 ~ 7e-7 Ha of
    # difference from the exact derivative (the exact gradient is pinned
    # against the oracle's density and force above)
    ok = torch.autograd.gradcheck(lambda tt: f(tt[0]), (t,), eps=1e-4, atol=7e-7, rtol=1e-6,
                                  nondet_tol=0.0)
    assert ok
