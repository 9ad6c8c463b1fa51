"""ORACLE -- test infrastructure only.

Restatement (numpy, complex128) of the reference's complex-wavefunction
branch for one k point of an orthorhombic cell:

* op_kinetic_sol with zcoef_kin (kinetic_sol_module.f90:72-125, coefficients
  kinetic_module.f90:118-149): c0 + |k|^2/2 on the diagonal,
  zcoef(+m) = coef_lap - i coef_nabk on psi(x + m), its conjugate on psi(x - m);
* op_localpot (localpot_module.f90:27-45);
* op_nonlocal with Bloch-phased uVk (ps_nloc2_op_module.f90:161-175, 255-263,
  304-313): c = iuV dV sum conj(uVk) psi(JJP), h(JJP) += uVk c;
* precond_hprsdft on complex gradients (cgpc_hprsdft_module.f90:126-263: one
  CG with Hermitian inner products, the real 7-point A of precond_cg_mat2);
* conjugate_gradient_1, complex branch (cg_module.f90:97-430 with the zhegv
  2x2 problem of :299-318: the band-choice rule, the phase fix making the
  first coefficient real and positive, the reject rule rb/res > 1e8);
* Gram-Schmidt (the orthonormal Q with a positive real diagonal of R, as
  gram_schmidt_t returns it) and subspace_diag_la's Hermitian Rayleigh-Ritz
  (zheev, ascending).
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import eigh


class KpointOracle:
    def __init__(self, system, proj, const_k2: float, coef_nabk: np.ndarray, vloc=None,
                 Ncg: int = 3, mloop: int = 3):
        g = system.grid
        self.N1, self.N2, self.N3 = g.N
        self.ML = g.ML
        self.dV = g.dV
        self.h = np.array(g.h)
        self.Md = system.Md
        self.coef_lap0 = system.coef_lap0
        self.coef_lap = np.asarray(system.coef_lap)          # (3, Md)
        self.coef_nabk = np.asarray(coef_nabk)               # (3, Md)
        self.const_k2 = const_k2
        self.vloc = np.asarray(system.vloc if vloc is None else vloc, dtype=np.float64)
        self.proj = proj
        self.Ncg, self.mloop = Ncg, mloop

    # --- operators ----------------------------------------------------------------
    def _grid(self, psi):
        return psi.reshape(psi.shape[0], self.N3, self.N2, self.N1)

    def op_kinetic(self, psi):
        psi = np.atleast_2d(np.asarray(psi, dtype=np.complex128))
        p = self._grid(psi)
        out = (self.coef_lap0 + self.const_k2) * p
        axis = {0: 3, 1: 2, 2: 1}                            # x, y, z axes of (nb, N3, N2, N1)
        for d in range(3):
            for m in range(1, self.Md + 1):
                zc = self.coef_lap[d, m - 1] - 1j * self.coef_nabk[d, m - 1]
                out = out + zc * np.roll(p, -m, axis=axis[d]) + np.conj(zc) * np.roll(p, m, axis=axis[d])
        return out.reshape(psi.shape)

    def op_localpot(self, psi):
        return np.atleast_2d(psi) * self.vloc[None, :]

    def op_nonlocal(self, psi):
        psi = np.atleast_2d(np.asarray(psi, dtype=np.complex128))
        pr = self.proj
        out = np.zeros_like(psi)
        for l in range(pr.nzlma):
            n = pr.MJJ[l]
            if n == 0:
                continue
            j = pr.JJP[l, :n] - 1
            u = pr.uVk[l, :n]
            c = pr.iuV[l] * self.dV * (psi[:, j] @ np.conj(u))
            out[:, j] += c[:, None] * u[None, :]
        return out

    def hamiltonian(self, psi):
        return self.op_kinetic(psi) + self.op_localpot(psi) + self.op_nonlocal(psi)

    # --- preconditioner -------------------------------------------------------------
    def _pc_apply(self, g):
        p = g.reshape(self.N3, self.N2, self.N1)
        h1, h2, h3 = self.h
        c = 1.0 / h1 ** 2 + 1.0 / h2 ** 2 + 1.0 / h3 ** 2
        c1, c2, c3 = -0.5 / h1 ** 2, -0.5 / h2 ** 2, -0.5 / h3 ** 2
        f = (c * p + c1 * (np.roll(p, -1, 2) + np.roll(p, 1, 2))
             + c2 * (np.roll(p, -1, 1) + np.roll(p, 1, 1))
             + c3 * (np.roll(p, -1, 0) + np.roll(p, 1, 0)))
        return f.ravel()

    def precond(self, gk):
        Pgk = gk.copy()
        if self.mloop <= 0:
            return Pgk
        f = self._pc_apply(gk)
        rk = gk - f
        pk = rk.copy()
        rr1 = float(np.sum(np.abs(rk) ** 2))
        if rr1 < 1.0e-24:
            return Pgk
        for iloop in range(1, self.mloop + 1):
            f = self._pc_apply(pk)
            pAp = float(np.real(np.vdot(pk, f)))
            rr0 = rr1
            a = rr0 / pAp
            Pgk = Pgk + a * pk
            if iloop == self.mloop:
                break
            rk = rk - a * f
            rr1 = float(np.sum(np.abs(rk) ** 2))
            pk = rk + (rr1 / rr0) * pk
        return Pgk

    # --- eigensolver pieces -----------------------------------------------------------
    @staticmethod
    def hegv2(A, B):
        """zhegv(1,'V','U') of a 2x2 Hermitian pair: A z = w B z, w ascending,
        z^H B z = 1."""
        w, Z = eigh(A, B)
        return w, Z

    def conjugate_gradient(self, psi, counts=False):
        x = np.array(psi, dtype=np.complex128)
        MB = x.shape[0]
        dV, c1 = self.dV, 2.0
        esp, res = np.zeros(MB), np.zeros(MB)
        nrej = 0
        for n in range(MB):
            xn = x[n]
            hx = self.hamiltonian(xn)[0]
            E1 = 1.0e10
            E = float(np.real(np.vdot(xn, hx))) * dV
            gk = -c1 * (hx - E * xn)
            rb = float(np.sum(np.abs(gk) ** 2)) * dV
            gkgk = 0.0
            pk = None
            for icg in range(1, self.Ncg + 2):
                Pgk = gk.copy()
                res[n] = rb / c1 ** 2
                if rb < 0.0 or abs(E - E1) < 1.0e-15 or icg == self.Ncg + 1:
                    break
                Pgk = self.precond(gk)
                rb = float(np.real(np.vdot(Pgk, gk))) * dV
                if icg == 1:
                    pk = Pgk.copy()
                else:
                    pk = Pgk + (rb / gkgk) * pk
                gkgk = rb
                hp = self.hamiltonian(pk)[0]
                v1 = np.real(np.vdot(xn, xn)) * dV
                v2 = np.vdot(pk, xn) * dV
                v3 = np.real(np.vdot(pk, pk)) * dV
                v4 = np.real(np.vdot(xn, hx)) * dV
                v5 = np.vdot(pk, hx) * dV
                v6 = np.real(np.vdot(pk, hp)) * dV
                B = np.array([[v1, np.conj(v2)], [v2, v3]], dtype=np.complex128)
                A = np.array([[v4, np.conj(v5)], [v5, v6]], dtype=np.complex128)
                W, Z = self.hegv2(A, B)
                z = Z[:, 0].copy()
                w = W[0]
                if abs(W[0] - E) > 1.0e-1 and abs(W[1] - E) <= 1.0e-1:
                    z = Z[:, 1].copy()
                    w = W[1]
                z = z * np.conj(z[0]) / abs(z[0])
                u1, u2 = z[0], z[1]
                E1, E = E, float(w)
                hx = u1 * hx + u2 * hp
                gk = -c1 * (hx - w * (u1 * xn + u2 * pk))
                rb = float(np.sum(np.abs(gk) ** 2)) * dV
                if rb / res[n] > 1.0e8:
                    E = E1
                    nrej += 1
                    continue
                xn = u1 * xn + u2 * pk
            x[n] = xn
            esp[n] = E
        if counts:
            return x, esp, res, nrej
        return x, esp, res

    def gram_schmidt(self, psi):
        """Orthonormal rows with dV sum conj(q_m) q_n = delta (classical GS
        with reorthogonalisation: Q of X = R^T Q, R upper with a positive real
        diagonal)."""
        X = np.array(psi, dtype=np.complex128)
        for n in range(X.shape[0]):
            for _ in range(2):
                if n:
                    c = (np.conj(X[:n]) @ X[n]) * self.dV
                    X[n] = X[n] - c @ X[:n]
            X[n] /= np.sqrt(np.real(np.vdot(X[n], X[n])) * self.dV)
        return X

    def subspace_diag(self, psi):
        U = np.asarray(psi, dtype=np.complex128)
        HU = self.hamiltonian(U)
        C = np.conj(U) @ HU.T
        Hsub = 0.5 * self.dV * (C + np.conj(C.T))
        w, V = np.linalg.eigh(Hsub)
        return np.ascontiguousarray(V.T @ U), w, Hsub

    def sweep(self, psi):
        x, _, res = self.conjugate_gradient(psi)
        x = self.gram_schmidt(x)
        x, esp, _ = self.subspace_diag(x)
        return x, esp, res
