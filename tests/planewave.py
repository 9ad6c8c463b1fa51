"""Plane-wave model of one CG step, for pinning the CG rules of the oracle.

On a periodic grid with a constant local potential and no projectors, real
plane waves ``cos(2 pi n.r / L)`` are eigenvectors of the order-2Md kinetic
stencil (eigenvalue in closed form from the FD weights) and of the 7-point
preconditioner matrix A of ``cgpc_hprsdft_module.f90`` (eigenvalue
``sum_d (1 - cos(2 pi n_d / N_d)) / h_d^2``).  A wavefunction made of a few
plane waves is then a short coefficient vector, and one step of
``conjugate_gradient_1`` (``cg_module.f90:186-366``) is small dense algebra
in that basis, computed here from other formulations than the oracle's:

* the preconditioned gradient from the Galerkin (A-norm optimal) solution
  on the Krylov space ``g + K_m(A, g - A g)``, which the m CG steps of
  ``precond_hprsdft`` reach in exact arithmetic;
* the 2x2 Rayleigh-Ritz problem from ``scipy.linalg.eigh(A, B)``;
* the band choice (``cg_module.f90:319-323``), the sign fix (``:325-329``)
  and the reject test ``rb/res > 1e8`` (``:361-364``) as stated in the
  reference.

The tests use it to build cases where each rule must fire (or must not)
and check the oracle, and the GPU, against the model's outcome.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sl

from rsdft_amd.system import Grid, Projectors, System, kinetic_coefficients


def no_projectors() -> Projectors:
    z = np.zeros(0, np.int32)
    return Projectors(0, z, np.zeros((0, 1), np.int32), np.zeros((0, 1)), z, z, z, z, z,
                      np.zeros((0, 1), np.int32), np.zeros((0, 1, 3)), 1.0)


class PlaneWaveCell:
    """Orthorhombic N1 x N2 x N3 grid of spacing h, V_loc = 0, no projectors."""

    def __init__(self, N=(20, 24, 28), h=0.5, Md=6, nband=3):
        self.N = tuple(N)
        self.grid = Grid(self.N, tuple(n * h for n in self.N))
        self.Md = Md
        self.c0, self.cl = kinetic_coefficients(Md, self.grid.h)
        self.system = System("planewave", self.grid, Md, np.zeros((0, 3), np.int64),
                             np.zeros((0, 3)), np.zeros((0, 4)), np.zeros(self.grid.ML),
                             no_projectors(), nband, np.zeros(nband), 0, self.c0, self.cl)
        N1, N2, N3 = self.N
        i3, i2, i1 = np.meshgrid(np.arange(N3), np.arange(N2), np.arange(N1), indexing="ij")
        self._frac = (i1 / N1, i2 / N2, i3 / N3)
        L = self.grid.L
        self.volume = L[0] * L[1] * L[2]

    def lam(self, n) -> float:
        """Eigenvalue of the kinetic stencil on cos(2 pi n.r/L):
        coef_lap0 + 2 sum_{d,m} coef_lap(d,m) cos(2 pi n_d m / N_d)."""
        t = self.c0
        for d in range(3):
            for m in range(1, self.Md + 1):
                t += 2.0 * self.cl[d, m - 1] * math.cos(2.0 * math.pi * n[d] * m / self.N[d])
        return t

    def mu(self, n) -> float:
        """Eigenvalue of the 7-point preconditioner matrix A on the same wave."""
        h = self.grid.h
        return sum((1.0 - math.cos(2.0 * math.pi * n[d] / self.N[d])) / h[d] ** 2
                   for d in range(3))

    def wave(self, n) -> np.ndarray:
        """cos(2 pi n.r/L) normalised so that sum v^2 dV = 1 (n != 0, no Nyquist)."""
        f = self._frac
        ph = 2.0 * math.pi * (n[0] * f[0] + n[1] * f[1] + n[2] * f[2])
        return (np.cos(ph) * math.sqrt(2.0 / self.volume)).ravel()

    def vector(self, basis, coef) -> np.ndarray:
        return sum(c * self.wave(n) for n, c in zip(basis, coef))


def galerkin_precond(mu: np.ndarray, g: np.ndarray, m: int) -> np.ndarray:
    """x in g + K_m(A, r0), r0 = g - A g, minimising the A-norm error of A x = g
    (A = diag(mu)); the result of m exact CG steps started from x0 = g."""
    r0 = g - mu * g
    K = np.stack([mu ** j * r0 for j in range(m)], axis=1)
    Q, _ = np.linalg.qr(K)
    c = np.linalg.solve(Q.T @ (mu[:, None] * Q), Q.T @ r0)
    return g + Q @ c


def cg_step_model(cell: PlaneWaveCell, basis, cx, mloop: int = 3) -> dict:
    """One step of conjugate_gradient_1 (Ncg = 1) from x = sum cx_k wave(basis_k),
    |x| = 1.  Returns E0, the Ritz pair values W, whether the band choice
    swapped, the reject ratio rb_new / res and the resulting (E, coefficients)."""
    lam = np.array([cell.lam(n) for n in basis])
    mu = np.array([cell.mu(n) for n in basis])
    cx = np.asarray(cx, dtype=np.float64)
    E0 = float(cx @ (lam * cx))
    g = -2.0 * (lam * cx - E0 * cx)
    rb = float(g @ g)
    p = galerkin_precond(mu, g, mloop)
    B = np.array([[cx @ cx, p @ cx], [p @ cx, p @ p]])
    A = np.array([[cx @ (lam * cx), p @ (lam * cx)], [p @ (lam * cx), p @ (lam * p)]])
    W, Z = sl.eigh(A, B)
    swap = abs(W[0] - E0) > 1.0e-1 and abs(W[1] - E0) <= 1.0e-1
    k = 1 if swap else 0
    u = Z[:, k].copy()
    if u[0] < 0.0:
        u = -u
    cy = u[0] * cx + u[1] * p
    gn = -2.0 * (lam * cy - W[k] * cy)
    ratio = float(gn @ gn) / (rb / 4.0)
    rejected = ratio > 1.0e8
    return {"E0": E0, "W": W, "swap": swap, "ratio": ratio, "rejected": rejected,
            "E": E0 if rejected else float(W[k]), "coef": cx if rejected else cy}


# The three bands of the pinning cases (Ncg = 1, mloop = 3):
#  0  x mostly the excited wave (2,0,0) with lower waves mixed in: the lower
#     Ritz value lies > 0.1 Ha below E and the upper one within 0.1 Ha, so the
#     band-choice rule must keep the upper pair (no jump to a lower state);
#  1  x = (1,0,0) plus 1e-2 of four other waves: the step is taken
#     (rb_new / res ~ 1e4, between 1e2 and 1e8);
#  2  the same with 1e-6: rb_new / res ~ 1e12 > 1e8, the step must be rejected
#     (E back to E0, x unchanged).
def pinning_cases():
    b0 = [(2, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 1, 1)]
    c0 = np.array([1.0, 0.1, 0.08, 0.06, 0.05])
    cases = [(b0, c0 / np.linalg.norm(c0))]
    for d in (1.0e-2, 1.0e-6):
        b = [(1, 0, 0), (0, 0, 1), (0, 1, 0), (0, 2, 0), (1, 1, 0)]
        c = np.array([1.0, d, 0.8 * d, 0.3 * d, 0.2 * d])
        cases.append((b, c / np.linalg.norm(c)))
    return cases
