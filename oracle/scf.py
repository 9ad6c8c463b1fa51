"""ORACLE -- test infrastructure only.

CPU FP64 restatement of one RSDFT self-consistent-field iteration around the
fixed-potential sweep (scf_module.f90:146-366, Nspin = 1, Gamma point, real
wavefunctions, NCPP): the FFT Hartree solve (hartree_sol_module.f90:19-122
with the G sphere of ggrid_module.f90:55-95,199-217 and the 1/ML-normalised
forward transform of rsdft_fft_module.f90:18-92), the Perdew-Zunger 81 LDA
(xc_ldapz81_module.f90:17-210), the Methfessel-Paxton occupations
(fermi_mp_module.f90:96-307), the total and Harris energies
(total_energy_module.f90:56-339) and the density mixing of imix = 10
(mixing_module.f90:69-299, pulay_r2_mixing :452-539) with normalize_density
(density_module.f90:100-111).  The eigensolver steps come from
``oracle.oracle.Oracle``.

Only ``tests/`` and ``bench.py``'s CPU-baseline leg use this module, always as
the checker, never as the measured product.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


# ---------------------------------------------------------------------------
# Hartree potential
# ---------------------------------------------------------------------------
def g_sphere(grid):
    """G^2 on the FFT grid (numpy order, array (N3, N2, N1)) and the mask of
    the reciprocal vectors calc_hartree_sol keeps: |i_d| <= (N_d-1)/2 and
    G^2 < Ecut - 1e-14 with Gcut = min_d |b_d| (2 NG_d + 1) / 2
    (CalcCutoff_Ggrid, ggrid_module.f90:55-95, orthorhombic cell)."""
    N1, N2, N3 = grid.N
    b = [2.0 * math.pi / L for L in grid.L]
    ng = [(n - 1) // 2 for n in grid.N]
    gcut = 0.5 * min(b[d] * (2 * ng[d] + 1) for d in range(3))
    ecut = gcut * gcut
    idx = [np.fft.fftfreq(n, 1.0 / n).astype(np.int64) for n in (N1, N2, N3)]
    i1, i2, i3 = idx
    g2 = ((b[2] * i3)[:, None, None] ** 2 + (b[1] * i2)[None, :, None] ** 2
          + (b[0] * i1)[None, None, :] ** 2)
    inside = ((np.abs(i3) <= ng[2])[:, None, None] & (np.abs(i2) <= ng[1])[None, :, None]
              & (np.abs(i1) <= ng[0])[None, None, :])
    mask = inside & (g2 < ecut - 1e-14) & (g2 != 0.0)
    return g2, mask


def hartree(rho, grid):
    """calc_hartree_sol: V_H(G) = 4 pi rho(G) / G^2 on the sphere, G = 0
    dropped; E_H = dV/2 sum V_H rho.  Returns (vh, E_H)."""
    N1, N2, N3 = grid.N
    r3 = np.asarray(rho, np.float64).reshape(N3, N2, N1)
    rg = np.fft.fftn(r3) / r3.size
    g2, mask = g_sphere(grid)
    vg = np.zeros_like(rg)
    vg[mask] = rg[mask] * (4.0 * math.pi) / g2[mask]
    vh = np.real(np.fft.ifftn(vg) * r3.size).ravel()
    eh = 0.5 * grid.dV * float(np.sum(vh * np.asarray(rho)))
    return vh, eh


# ---------------------------------------------------------------------------
# LDA exchange-correlation (Perdew-Zunger 81), one spin channel
# ---------------------------------------------------------------------------
_PZ = dict(gam=(-0.1423, -0.0843), bet1=(1.0529, 1.3981), bet2=(0.3334, 0.2611),
           A=(0.0311, 0.01555), B=(-0.048, -0.0269), C=(0.002, 0.0007), D=(-0.0116, -0.0048))


def xc_pz81(rho, dV):
    """calc_LDAPZ81 with MSP = 1: returns (vxc, Exc, Ex, Ec).
    Exchange: e_x = -3/4 (3 rho / pi)^(1/3), v_x = 4/3 e_x
    (calc_ldapz81_x); correlation: the rs >= 1 Pade and rs < 1 logarithmic
    forms of the unpolarised gas, rho_a = rho_b = rho / 2 so the
    polarisation function f vanishes (calc_ldapz81_c)."""
    rho = np.asarray(rho, np.float64)
    onethr = 1.0 / 3.0
    exd = -0.75 * (3.0 / math.pi * rho) ** onethr
    ex = float(np.sum(rho * exd)) * dV
    vx = (4.0 / 3.0) * exd
    trho = np.abs(0.5 * rho) + np.abs(0.5 * rho)
    vc = np.zeros_like(rho)
    ec_d = np.zeros_like(rho)
    pos = trho > 0.0
    rs = np.zeros_like(rho)
    rs[pos] = (3.0 / (4.0 * math.pi) / trho[pos]) ** onethr
    g, b1, b2 = _PZ["gam"][0], _PZ["bet1"][0], _PZ["bet2"][0]
    A, B, C, D = _PZ["A"][0], _PZ["B"][0], _PZ["C"][0], _PZ["D"][0]
    hi = pos & (rs >= 1.0)
    lo = pos & (rs < 1.0)
    r = rs[hi]
    sq = np.sqrt(r)
    e1 = g / (1.0 + b1 * sq + b2 * r)
    ec_d[hi] = e1
    vc[hi] = e1 + e1 * onethr * (0.5 * b1 / sq + b2) / (1.0 / r + b1 / sq + b2)
    r = rs[lo]
    ln = np.log(r)
    e1 = A * ln + B + C * r * ln + D * r
    ec_d[lo] = e1
    vc[lo] = e1 - onethr * (A + C * r * (1.0 + ln) + r * D)
    ec = float(np.sum(trho * ec_d)) * dV
    return vx + vc, ex + ec, ex, ec


# ---------------------------------------------------------------------------
# occupations (Methfessel-Paxton, order kinteg)
# ---------------------------------------------------------------------------
def _mp_factors(kinteg):
    f = [0.0] * (kinteg + 1)
    if kinteg > 0:
        f[1] = -1.0 / (4.0 * math.sqrt(math.pi))
        for n in range(2, kinteg + 1):
            f[n] = -f[n - 1] / (4.0 * n)
    return f


def mp_occupation(x, kinteg, fac):
    """ff0 (fermi_mp_module.f90:289-307)."""
    ff0 = 0.5 * (1.0 - erf(x))
    if kinteg <= 0:
        return ff0
    hp0, hp1 = 1.0, 2.0 * x
    ff = fac[1] * hp1
    for i in range(2, kinteg + 1):
        hp2 = 2.0 * x * hp1 - 2.0 * (2 * i - 3) * hp0
        hp3 = 2.0 * x * hp2 - 2.0 * (2 * i - 2) * hp1
        ff = ff + fac[i] * hp3
        hp0, hp1 = hp2, hp3
    return ff0 + ff * math.exp(-x * x)


def calc_fermi(esp, nelectron, ekbt=1e-5, kinteg=5):
    """calc_fermi (fermi_mp_module.f90:96-287), one spin, one k point, weight
    1: bisection of the Fermi level on the occupation count (eps = 0, at most
    1000 halvings, stop when the midpoint repeats), occ = 2 ff.  Returns
    (occ, efermi)."""
    esp = np.asarray(esp, np.float64)
    fac = _mp_factors(kinteg)
    ef1, ef2 = float(esp.min()), float(esp.max())
    if ef1 == ef2:
        ef1, ef2 = ef1 - 0.01, ef2 + 0.01
    ef2 = ef2 + min(ekbt * 1e2, 0.1)
    zne = nelectron
    occ = np.zeros_like(esp)
    ef0 = 1e10
    ef = 0.5 * (ef1 + ef2)
    converged = False
    for _ in range(1000):
        ef = 0.5 * (ef1 + ef2)
        if ef == ef0:
            converged = True
            break
        oc = 0.0
        for n in range(len(esp)):
            ff = mp_occupation((esp[n] - ef) / ekbt, kinteg, fac)
            oc += ff * 2.0
            occ[n] = ff
        if oc - zne > 0.0:
            ef2 = ef
        elif oc - zne < 0.0:
            ef1 = ef
        else:
            converged = True
            break
        ef0 = ef
    if not converged and abs(oc - zne) > 1e-10:
        raise RuntimeError("FERMI: the Fermi level did not converge")
    return 2.0 * occ, ef


# ---------------------------------------------------------------------------
# density mixing, imix = 10 (Pulay on rho), one spin
# ---------------------------------------------------------------------------
class Mixer:
    """init_mixing / perform_mixing / calc_sqerr_mixing for imix = 10.
    The reference keeps Xin/Xou complex; every value here is real (real
    inputs, a real symmetric A0 and zgesv of it), so real arrays hold the
    same numbers."""

    def __init__(self, rho, vloc, dV, mmix=4, beta=1.0):
        self.mmix, self.beta, self.dV = mmix, beta, dV
        ML = rho.size
        self.ML = ML
        self.Xold_rho = np.array(rho, np.float64)
        self.Xold_v = np.array(vloc, np.float64)
        self.Xin = np.zeros((mmix, ML))
        self.Xou = np.zeros((mmix, ML))
        self.Xin[mmix - 1] = self.Xold_rho
        self.count = 0
        self.dif0 = 0.0

    def sqerr(self, rho, v):
        """calc_sqerr: (sum (rho - rho_old)^2 / ML, sum (v - v_old)^2 / ML)."""
        return (float(np.sum((rho - self.Xold_rho) ** 2)) / self.ML,
                float(np.sum((v - self.Xold_v) ** 2)) / self.ML)

    def _pulay_r2(self, f, beta):
        m = self.mmix
        self.Xou[m - 1] = f
        self.count += 1
        m0 = min(self.count, m)
        if m == 1 or m0 < m:
            self.Xin[:m - 1] = self.Xin[1:].copy()
            self.Xou[:m - 1] = self.Xou[1:].copy()
            self.Xin[m - 1] = self.Xin[m - 1] + beta * (self.Xou[m - 1] - self.Xin[m - 1])
            return self.Xin[m - 1].copy()
        R = self.Xou - self.Xin
        A = R @ R.T
        b = np.linalg.solve(A, np.ones(m))
        b = b * (1.0 / np.sum(b))
        X = b @ self.Xin
        Y = b @ self.Xou
        self.Xin[:m - 1] = self.Xin[1:].copy()
        self.Xou[:m - 1] = self.Xou[1:].copy()
        self.Xin[m - 1] = X + beta * (Y - X)
        return self.Xin[m - 1].copy()

    def mix(self, rho_out, v_out):
        """perform_mixing: returns the mixed density (|h|, chk_chrg) and
        records (rho, v) as the new old pair.  beta >= 1 allows up to ten
        retries with beta halved while the change ratio exceeds 0.5."""
        f = np.array(rho_out, np.float64)
        h_old = self.Xold_rho
        bak = (self.Xin.copy(), self.Xou.copy(), self.count)
        beta = self.beta
        beta_min, dif_min = beta, 1e10
        max_loop = 11 if self.beta >= 1.0 else 1
        for loop in range(1, max_loop + 1):
            h = self._pulay_r2(f, beta)
            dif1 = float(np.sum((h - h_old) ** 2)) / self.ML
            diff = dif1 / self.dif0 if self.dif0 != 0.0 else 0.0
            if dif1 < dif_min:
                beta_min, dif_min = beta, dif1
            if loop < max_loop:
                if diff > 0.5:
                    self.Xin, self.Xou, self.count = bak[0].copy(), bak[1].copy(), bak[2]
                    beta = beta * 0.5
                else:
                    break
                if loop == max_loop - 1:
                    beta = beta_min
        self.dif0 = dif1
        h = np.abs(h)
        self.Xold_rho = h.copy()
        self.Xold_v = np.array(v_out, np.float64)
        return h


def normalize_density(rho, nelectron, dV):
    return rho * (nelectron / (float(np.sum(rho)) * dV))


# ---------------------------------------------------------------------------
# one SCF iteration
# ---------------------------------------------------------------------------
class Scf:
    """scf loop body (scf_module.f90:146-366) for a system whose ionic local
    potential is ``vion``: Vloc = Vion + V_H[rho] + V_xc[rho].  Energies
    follow calc_total_energy with Eewald = 0 and const_ps_local = 0 (the
    synthetic ionic potential carries no point charges)."""

    def __init__(self, system, rho0, vion=None, nelectron=None, mmix=4, beta=1.0, ekbt=1e-5,
                 kinteg=5, Ncg=3, mloop=3):
        from oracle.oracle import Oracle
        self.sys = system
        self.grid = system.grid
        self.dV = system.grid.dV
        self.vion = np.array(system.vloc if vion is None else vion, np.float64)
        self.nel = float(np.sum(system.occ)) if nelectron is None else float(nelectron)
        self.ekbt, self.kinteg = ekbt, kinteg
        self.Ncg, self.mloop = Ncg, mloop
        self.rho = np.array(rho0, np.float64)
        self.vh, self.eh = hartree(self.rho, self.grid)
        self.vxc, self.exc, _, _ = xc_pz81(self.rho, self.dV)
        self.vloc = self.vion + self.vh + self.vxc
        self.mixer = Mixer(self.rho, self.vloc, self.dV, mmix, beta)
        self._Oracle = Oracle

    def oracle(self):
        return self._Oracle(self.sys, vloc=self.vloc, Ncg=self.Ncg, mloop=self.mloop)

    def iterate(self, psi):
        """One iteration from psi; returns (psi, esp, occ, info dict)."""
        o = self.oracle()
        x, _, _ = o.subspace_diag(psi)
        x, _, res = o.conjugate_gradient(x)
        x = o.gram_schmidt(x)
        hx = o.hamiltonian(x)
        esp = np.einsum("ij,ij->i", x, hx) * self.dV            # esp_calc
        occ, ef = calc_fermi(esp, self.nel, self.ekbt, self.kinteg)
        dV = self.dV
        # Harris energy from the input density (calc_with_rhoIN_total_energy)
        eeig = float(np.sum(occ * esp))
        eloc_in = float(np.sum(self.vloc * self.rho)) * dV
        eion_in = float(np.sum(self.vion * self.rho)) * dV
        ehwf = eeig - eloc_in + self.eh + self.exc + eion_in
        # output density and its energies
        rho = o.density(x, occ)
        vh, eh = hartree(rho, self.grid)
        vxc, exc, _, _ = xc_pz81(rho, dV)
        eloc = float(np.sum(rho * self.vloc)) * dV
        eion = float(np.sum(rho * self.vion)) * dV
        etot = eeig - eloc + eh + exc + eion
        vout = self.vion + vh + vxc
        sq = self.mixer.sqerr(rho, vout)
        # mixing, then the potential of the mixed density
        rho_in = self.mixer.mix(rho, vout)
        rho_in = normalize_density(rho_in, self.nel, dV)
        self.rho = rho_in
        self.vh, self.eh = hartree(rho_in, self.grid)
        self.vxc, self.exc, _, _ = xc_pz81(rho_in, dV)
        self.vloc = self.vion + self.vh + self.vxc
        info = dict(etot=etot, ehwf=ehwf, eeig=eeig, eloc=eloc, eion=eion, eh=eh, exc=exc,
                    efermi=ef, sqerr_rho=sq[0], sqerr_v=sq[1], res=res)
        return x, esp, occ, info
