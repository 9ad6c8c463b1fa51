"""ORACLE -- test infrastructure only.

CPU FP64 restatement of RSDFT's Kohn-Sham sweep at fixed potential.  The
grid-point loops (hamiltonian, preconditioner, CG, density, force) are plain C
in ``rsdft_oracle.c``; the dense parts that the reference hands to BLAS/LAPACK
(the dgemm/dgemv blocks of Gram-Schmidt, dsyr2k, DSYEV, the rotation dgemm)
are numpy calls here, in the reference's block order.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg use this module, always as the checker, never as the measured product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")


def build(force: bool = False) -> str:
    """Compile the C restatement with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class _Ham(ctypes.Structure):
    _fields_ = [
        ("N1", ctypes.c_int), ("N2", ctypes.c_int), ("N3", ctypes.c_int), ("Md", ctypes.c_int),
        ("h", ctypes.c_double * 3), ("dV", ctypes.c_double), ("coef_lap0", ctypes.c_double),
        ("coef_lap", ctypes.c_void_p), ("vloc", ctypes.c_void_p),
        ("nzlma", ctypes.c_int), ("maxmjj", ctypes.c_int),
        ("MJJ", ctypes.c_void_p), ("JJP", ctypes.c_void_p), ("uVk", ctypes.c_void_p),
        ("iuV", ctypes.c_void_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(build())
        P, I, L, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        _lib.orc_fd_coef.argtypes = [I, P]
        for name in ("orc_op_kinetic", "orc_op_localpot", "orc_op_nonlocal", "orc_hamiltonian"):
            getattr(_lib, name).argtypes = [P, I, P, P]
        _lib.orc_precond_apply.argtypes = [P, P, P]
        _lib.orc_precond.argtypes = [P, I, P, P, P]
        _lib.orc_sygv2.argtypes = [P, P, P, P]
        _lib.orc_cg.argtypes = [P, I, I, I, P, P, P, P, P]
        _lib.orc_density.argtypes = [L, I, P, P, P]
        _lib.orc_force_nloc.argtypes = [P, I, I, P, P, P, P, I, P, P, P]
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def fd_coef(Md: int) -> np.ndarray:
    """lap(-Md..Md) from get_coef_lapla_fd (fd_module.f90:18-54)."""
    out = np.zeros(2 * Md + 1)
    lib().orc_fd_coef(Md, _p(out))
    return out


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """Hamiltonian, CG, Gram-Schmidt and subspace diagonalisation for one system.

    ``system`` is a :class:`rsdft_amd.system.System`; optional ``vloc`` and
    ``proj`` override its potential or projector tables.
    """

    def __init__(self, system, vloc=None, proj=None, Ncg: int = 3, mloop: int = 3):
        g = system.grid
        self.sys = system
        self.grid = g
        self.ML = g.ML
        self.dV = g.dV
        self.Ncg, self.mloop = Ncg, mloop
        self._coef = _f64(np.asarray(system.coef_lap).T.ravel())   # Fortran (3, Md)
        self._vloc = _f64(system.vloc if vloc is None else vloc)
        pr = system.proj if proj is None else proj
        self.proj = pr
        self._MJJ = np.ascontiguousarray(pr.MJJ, np.int32)
        self._JJP = np.ascontiguousarray(pr.JJP, np.int32)
        self._uVk = _f64(pr.uVk)
        self._iuV = np.ascontiguousarray(pr.iuV, np.int32)
        self._amap = np.ascontiguousarray(pr.amap, np.int32)
        self._MJJ_MAP = np.ascontiguousarray(pr.MJJ_MAP, np.int32)
        self._JJ_MAP = np.ascontiguousarray(pr.JJ_MAP, np.int32)
        self._duVdR = _f64(pr.duVdR)
        H = _Ham()
        H.N1, H.N2, H.N3 = g.N
        H.Md = system.Md
        H.h = (ctypes.c_double * 3)(*g.h)
        H.dV = g.dV
        H.coef_lap0 = system.coef_lap0
        H.coef_lap = self._coef.ctypes.data
        H.vloc = self._vloc.ctypes.data
        H.nzlma = pr.nzlma
        H.maxmjj = pr.MAXMJJ
        H.MJJ = self._MJJ.ctypes.data
        H.JJP = self._JJP.ctypes.data
        H.uVk = self._uVk.ctypes.data
        H.iuV = self._iuV.ctypes.data
        self._H = H
        self._Hp = ctypes.cast(ctypes.pointer(H), ctypes.c_void_p)

    # --- operators -----------------------------------------------------------
    def _op(self, name, psi, out=None, accumulate=False):
        psi = _f64(np.atleast_2d(psi))
        if out is None:
            out = np.zeros_like(psi)
        getattr(lib(), name)(self._Hp, psi.shape[0], _p(psi), _p(out))
        return out

    def hamiltonian(self, psi):
        return self._op("orc_hamiltonian", psi)

    def op_kinetic(self, psi):
        return self._op("orc_op_kinetic", psi)

    def op_localpot(self, psi, vf=None):
        vf = np.zeros_like(_f64(np.atleast_2d(psi))) if vf is None else vf
        return self._op("orc_op_localpot", psi, out=vf)

    def op_nonlocal(self, psi, hf=None):
        hf = np.zeros_like(_f64(np.atleast_2d(psi))) if hf is None else hf
        return self._op("orc_op_nonlocal", psi, out=hf)

    def precond_apply(self, g):
        g = _f64(g)
        f = np.zeros_like(g)
        lib().orc_precond_apply(self._Hp, _p(g), _p(f))
        return f

    def precond(self, g):
        g = _f64(g)
        Pg = g.copy()
        work = np.zeros(3 * self.ML)
        lib().orc_precond(self._Hp, self.mloop, _p(g), _p(Pg), _p(work))
        return Pg

    # --- eigensolver pieces ----------------------------------------------------
    def conjugate_gradient(self, psi, counts=False):
        """conjugate_gradient_1 (cg_module.f90:97-430); returns (psi, esp, res)."""
        x = _f64(psi).copy()
        MB = x.shape[0]
        esp, res = np.zeros(MB), np.zeros(MB)
        work = np.zeros(9 * self.ML)
        cnt = np.zeros(2, np.int32)
        lib().orc_cg(self._Hp, MB, self.Ncg, self.mloop, _p(x), _p(esp), _p(res), _p(work), _p(cnt))
        if counts:
            return x, esp, res, (int(cnt[0]), int(cnt[1]))
        return x, esp, res

    def gram_schmidt(self, psi, NBLK: int = 0, NBLK1: int = 0):
        """gram_schmidt_t (gram_schmidt_t_module.f90:84-316), one band rank."""
        U = _f64(psi).copy()
        MB = U.shape[0]
        NBLK = NBLK or MB
        NBLK1 = NBLK1 or 4
        ncycle = (MB - 1) // NBLK + 1
        for k1 in range(1, ncycle + 1):
            ns = NBLK * (k1 - 1) + 1
            ne = min(ns + NBLK - 1, MB)
            _gs_sub(U, ns, ne, ns, ne, NBLK, NBLK1, self.dV)
            if ns <= MB - NBLK:
                for nbss in range(k1 + 1, ncycle + 1):
                    ms = NBLK * (nbss - 1) + 1
                    me = min(ms + NBLK - 1, MB)
                    if ms <= me:
                        _gs_sub(U, ms, me, ns, ne, NBLK, NBLK1, self.dV)
        return U

    def subspace_diag(self, psi):
        """subspace_diag_la (subspace_diag_la_module.f90:85-206).

        Hsub = 0.5 dV (Psi^T H Psi + (H Psi)^T Psi) (dsyr2k 'U','T'), DSYEV
        ascending, Psi <- Psi V.  Returns (psi, esp, Hsub).
        """
        U = _f64(psi)
        HU = self.hamiltonian(U)
        C = U @ HU.T
        Hsub = 0.5 * self.dV * (C + C.T)
        w, V = np.linalg.eigh(Hsub, UPLO="U")
        return np.ascontiguousarray(V.T @ U), w, Hsub

    def sweep(self, psi):
        """One calc_sweep iteration body (sweep_module.f90:147-158)."""
        x, _, res = self.conjugate_gradient(psi)
        x = self.gram_schmidt(x)
        x, esp, _ = self.subspace_diag(x)
        return x, esp, res

    # --- adjoint references ----------------------------------------------------
    def density(self, psi, occ):
        """calc_density (density_module.f90:129-136): rho = sum occ psi^2."""
        psi = _f64(np.atleast_2d(psi))
        occ = _f64(occ)
        rho = np.zeros(self.ML)
        lib().orc_density(self.ML, psi.shape[0], _p(occ), _p(psi), _p(rho))
        return rho

    def force_nloc(self, psi, occ):
        """calc_force_ps_nloc2 (ps_nloc2_module.f90:1601-1773), shape (natom, 3)."""
        psi = _f64(np.atleast_2d(psi))
        occ = _f64(occ)
        nat = self.sys.natom
        F = np.zeros(3 * nat)
        pr = self.proj
        lib().orc_force_nloc(self._Hp, psi.shape[0], nat, _p(occ), _p(psi), _p(self._amap),
                             _p(self._MJJ_MAP), pr.MMJJ, _p(self._JJ_MAP), _p(self._duVdR), _p(F))
        return F.reshape(nat, 3)

    def band_energy(self, psi, occ):
        """E = sum_n occ_n <psi_n|H|psi_n> dV."""
        psi = _f64(np.atleast_2d(psi))
        Hpsi = self.hamiltonian(psi)
        return float(np.sum(np.asarray(occ) * np.einsum("ij,ij->i", psi, Hpsi)) * self.dV)


def _gs_sub(U, mm1, mm2, nn1, nn2, MBLK, NBLK1, dV):
    """gram_schmidt_sub (gram_schmidt_t_module.f90:161-316); 1-based band
    indices, U[b] is band b+1.  dgemm/dgemv become numpy products."""
    for ms in range(mm1, mm2 + 1, MBLK):
        me = min(ms + MBLK - 1, mm2)
        mm = me - ms + 1
        for ns in range(nn1, nn2 + 1, MBLK):
            ne = min(ns + MBLK - 1, nn2)
            ne = min(ne, me - 1)
            nn = ne - ns + 1
            if nn <= 0:
                continue
            if ms >= ne + 1:
                utmp2 = -dV * (U[ns - 1:ne] @ U[ms - 1:me].T)          # (nn, mm)
                U[ms - 1:me] += utmp2.T @ U[ns - 1:ne]
                if ms == ne + 1:
                    c = float(np.dot(U[ms - 1], U[ms - 1]))
                    U[ms - 1] *= 1.0 / np.sqrt(c * dV)
            elif mm <= NBLK1:
                for m in range(ms, me + 1):
                    n = min(m - 1, ne)
                    if n - ns + 1 > 0:
                        utmp = -dV * (U[ns - 1:n] @ U[m - 1])
                        U[m - 1] += utmp @ U[ns - 1:n]
                    if m == 1 or (n == m - 1 and m != ns):
                        c = float(np.dot(U[m - 1], U[m - 1]))
                        U[m - 1] *= 1.0 / np.sqrt(c * dV)
            else:
                MBLKH = max(MBLK // 2, NBLK1)
                _gs_sub(U, ms, me, ns, ne, MBLKH, NBLK1, dV)
