"""Python host mirror of the reference's operator interface over the C-ABI.

``RsdftGpu`` loads ``librsdft_gpu.so`` (built in-tree by ``rsdft_amd.build``)
and exposes the reference's routines with their argument meaning:
``hamiltonian`` / ``op_kinetic`` / ``op_localpot`` / ``op_nonlocal``
(hamiltonian_module.f90:20), ``conjugate_gradient`` (cg_module.f90:54),
``gram_schmidt`` (gram_schmidt_module.f90:28), ``subspace_diag``
(subspace_diag_module.f90:17), ``sweep`` (sweep_module.f90:147-158),
``calc_density`` (density_module.f90:115) and ``calc_force_ps_nloc2``
(ps_nloc2_module.f90:1281).  Arrays are numpy, band-major ``(nb, ML)``.
A non-zero status from the library raises :class:`RsdftGpuError` carrying
its message, as the reference's ``stop`` would end the run.

There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import atexit
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsdft_gpu.so")

PREC_FP64 = 0
PREC_MIXED_FP32 = 1
PREC_MIXED_BF16 = 2

# every C-ABI entry point of include/rsdft_gpu.h
EXPORTS = (
    "rsdft_gpu_init", "rsdft_gpu_set_vloc", "rsdft_gpu_set_projectors",
    "rsdft_gpu_set_projector_derivatives", "rsdft_gpu_set_precision",
    "rsdft_gpu_set_solver_params", "rsdft_gpu_set_precond_mode", "rsdft_gpu_set_eigensolver",
    "rsdft_gpu_set_sweep_mode",
    "rsdft_gpu_hamiltonian", "rsdft_gpu_hamiltonian_noalloc", "rsdft_gpu_op_kinetic",
    "rsdft_gpu_op_localpot", "rsdft_gpu_op_nonlocal", "rsdft_gpu_wf_attach",
    "rsdft_gpu_wf_sync", "rsdft_gpu_wf_detach", "rsdft_gpu_conjugate_gradient",
    "rsdft_gpu_gram_schmidt", "rsdft_gpu_subspace_diag", "rsdft_gpu_sweep",
    "rsdft_gpu_calc_density", "rsdft_gpu_force_ps_nloc2", "rsdft_gpu_band_energy_grad",
    "rsdft_gpu_calc_hartree", "rsdft_gpu_calc_xc", "rsdft_gpu_calc_fermi", "rsdft_gpu_scf_init",
    "rsdft_gpu_scf_iterate", "rsdft_gpu_scf_get_fields",
    "rsdft_gpu_test_hpsi", "rsdft_gpu_test_gemm", "rsdft_gpu_test_eigh", "rsdft_gpu_get_timers", "rsdft_gpu_reset_timers",
    "rsdft_gpu_finalize", "rsdft_gpu_last_error",
    "rsdft_gpu_write_wf", "rsdft_gpu_read_wf", "rsdft_gpu_write_vrho", "rsdft_gpu_read_vrho",
    "rsdft_gpu_set_kpoint", "rsdft_gpu_zhamiltonian", "rsdft_gpu_zconjugate_gradient",
    "rsdft_gpu_zgram_schmidt", "rsdft_gpu_zsubspace_diag", "rsdft_gpu_zsweep",
)


class RsdftGpuError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RsdftGpuError(f"{path} is missing: run rsdft_amd.build.build() first")
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    sig = {
        "rsdft_gpu_init": [I, I, I, P, D, I, P, D],
        "rsdft_gpu_set_vloc": [P, I],
        "rsdft_gpu_set_projectors": [I, I, P, P, P, P],
        "rsdft_gpu_set_projector_derivatives": [I, P, I, P, P, P],
        "rsdft_gpu_set_precision": [I],
        "rsdft_gpu_set_solver_params": [I, I],
        "rsdft_gpu_set_precond_mode": [I],
        "rsdft_gpu_set_eigensolver": [I],
        "rsdft_gpu_set_sweep_mode": [I],
        "rsdft_gpu_hamiltonian": [I, I, P, P, I, I, I, I],
        "rsdft_gpu_hamiltonian_noalloc": [I, I, P, P, I, I, I, I],
        "rsdft_gpu_op_kinetic": [I, P, P, I, I, I, I],
        "rsdft_gpu_op_localpot": [I, I, I, P, P],
        "rsdft_gpu_op_nonlocal": [I, I, P, P, I, I, I, I],
        "rsdft_gpu_wf_attach": [P, I, I],
        "rsdft_gpu_wf_sync": [I],
        "rsdft_gpu_wf_detach": [],
        "rsdft_gpu_conjugate_gradient": [I, I, I, I, I, P, P, P],
        "rsdft_gpu_gram_schmidt": [P, I, I, I, I],
        "rsdft_gpu_subspace_diag": [P, P, I, I],
        "rsdft_gpu_sweep": [P, P, P, I, I, I],
        "rsdft_gpu_calc_density": [P, I, I, P, P],
        "rsdft_gpu_force_ps_nloc2": [P, I, I, P, P],
        "rsdft_gpu_band_energy_grad": [P, I, I, P, P, P, P],
        "rsdft_gpu_calc_hartree": [P, P, P],
        "rsdft_gpu_calc_xc": [P, P, P],
        "rsdft_gpu_calc_fermi": [P, I, D, D, I, P, P],
        "rsdft_gpu_scf_init": [P, P, D, I, D, D, I],
        "rsdft_gpu_scf_iterate": [P, I, I, P, P, P, P],
        "rsdft_gpu_scf_get_fields": [P, P, P, P],
        "rsdft_gpu_test_eigh": [P, I, I, P, P, P],
        "rsdft_gpu_test_hpsi": [I, P],
        "rsdft_gpu_test_gemm": [I, P],
        "rsdft_gpu_get_timers": [P, I],
        "rsdft_gpu_reset_timers": [],
        "rsdft_gpu_finalize": [],
        "rsdft_gpu_write_wf": [ctypes.c_char_p, I, I, I, P, I, I, I, I, I, P, P, P],
        "rsdft_gpu_read_wf": [ctypes.c_char_p, I, I, I, P, I, I, P, P],
        "rsdft_gpu_write_vrho": [ctypes.c_char_p, I, I, I, I, I, P, P, P, P, P, I, I, P, P, P, D],
        "rsdft_gpu_read_vrho": [ctypes.c_char_p, I, I, I, I, I, P, P, P, P, P],
        "rsdft_gpu_set_kpoint": [D, P, I, I, P, P, P, P],
        "rsdft_gpu_zhamiltonian": [I, I, P, P, I, I, I, I],
        "rsdft_gpu_zconjugate_gradient": [P, P, P, I, I],
        "rsdft_gpu_zgram_schmidt": [P, I, I],
        "rsdft_gpu_zsubspace_diag": [P, P, I, I],
        "rsdft_gpu_zsweep": [P, P, P, I, I, I],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = I
    lib.rsdft_gpu_last_error.argtypes = []
    lib.rsdft_gpu_last_error.restype = ctypes.c_char_p
    _lib = lib
    # release device memory, the stream and the rocBLAS handle while the HIP
    # runtime is still up (Python's exit handlers run before the C runtime's)
    atexit.register(lib.rsdft_gpu_finalize)
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _check(lib, name, r):
    if r != 0:
        msg = lib.rsdft_gpu_last_error().decode(errors="replace")
        raise RsdftGpuError(f"{name} -> {r}: {msg}")


def calc_fermi(esp, nelectron, ekbt=1e-5, kinteg=5):
    """calc_fermi with Methfessel-Paxton smearing (fermi_mp_module.f90:96-287):
    returns (occ, efermi)."""
    lib = load_library()
    e = _f64(esp)
    occ = np.zeros(e.size)
    ef = np.zeros(1)
    _check(lib, "calc_fermi", lib.rsdft_gpu_calc_fermi(_p(e), e.size, float(nelectron),
                                                     float(ekbt), int(kinteg), _p(occ), _p(ef)))
    return occ, float(ef[0])


def write_wf(path, N, unk, occ, OC=1, MBwr=(1, 0), aa=None, bb=None):
    """simple_wf_io_write (io_write_module.f90:19-264) of unk (MB, ML) on the
    periodic grid N = (N1, N2, N3)."""
    lib = load_library()
    u = unk if (unk.dtype == np.float64 and unk.flags.c_contiguous) else _f64(unk)
    occ = _f64(occ)
    a = None if aa is None else np.asfortranarray(aa, dtype=np.float64)
    b = None if bb is None else np.asfortranarray(bb, dtype=np.float64)
    _check(lib, "write_wf", lib.rsdft_gpu_write_wf(
        os.fsencode(path), N[0], N[1], N[2], _p(u), u.shape[1], u.shape[0], int(OC),
        int(MBwr[0]), int(MBwr[1]), _p(occ), _p(a), _p(b)))


def read_wf(path, N, MB, unk=None):
    """simple_wf_io_read (io_read_module.f90:15-521): returns (unk (MB, ML),
    occ (MB,), MB of the file); unk may be given (e.g. the attached array)."""
    lib = load_library()
    ML = N[0] * N[1] * N[2]
    if unk is None:
        unk = np.zeros((MB, ML))
    occ = np.zeros(MB)
    mbf = ctypes.c_int(0)
    _check(lib, "read_wf", lib.rsdft_gpu_read_wf(
        os.fsencode(path), N[0], N[1], N[2], _p(unk), ML, MB, _p(occ), ctypes.byref(mbf)))
    return unk, occ, mbf.value


def write_vrho(path, N, rho, vloc, vh, vxc, aa=None, aa_atom=None, ki_atom=None, zn_atom=None,
               efermi=0.0):
    """write_data's density/potential file (io_module.f90:96-200); rho, vloc,
    vxc (MSP, ML), vh (ML,)."""
    lib = load_library()
    rho, vloc, vxc = (np.atleast_2d(_f64(a)) for a in (rho, vloc, vxc))
    vh = _f64(vh)
    natom = 0 if aa_atom is None else len(ki_atom)
    nel = 0 if zn_atom is None else len(zn_atom)
    a = None if aa is None else np.asfortranarray(aa, dtype=np.float64)
    at = None if aa_atom is None else _f64(aa_atom)
    ki = None if ki_atom is None else _i32(ki_atom)
    zn = None if zn_atom is None else _i32(zn_atom)
    _check(lib, "write_vrho", lib.rsdft_gpu_write_vrho(
        os.fsencode(path), N[0], N[1], N[2], rho.shape[1], rho.shape[0], _p(rho), _p(vloc),
        _p(vh), _p(vxc), _p(a), nel, natom, _p(at), _p(ki), _p(zn), float(efermi)))


def read_vrho(path, N, MSP=1):
    """read_data's density/potential part (io_module.f90:286-409): returns
    (rho (MSP, ML), vloc (MSP, ML), vh (ML,), vxc (MSP, ML), efermi)."""
    lib = load_library()
    ML = N[0] * N[1] * N[2]
    rho, vloc, vxc = (np.zeros((MSP, ML)) for _ in range(3))
    vh = np.zeros(ML)
    ef = np.zeros(1)
    _check(lib, "read_vrho", lib.rsdft_gpu_read_vrho(
        os.fsencode(path), N[0], N[1], N[2], ML, MSP, _p(rho), _p(vloc), _p(vh), _p(vxc), _p(ef)))
    return rho, vloc, vh, vxc, float(ef[0])


class RsdftGpu:
    """One GPU context for one system (grid, Vloc, projector tables)."""

    def __init__(self, system, precision: int = PREC_FP64, ncg: int = 3, mloop: int = 3,
                 with_derivatives: bool = True, proj=None, vloc=None):
        self.lib = load_library()
        self.sys = system
        g = system.grid
        self.ML = g.ML
        self.dV = g.dV
        self._keep = []
        coef = _f64(np.asarray(system.coef_lap).T.ravel())       # Fortran coef_lap(3, Md)
        h = _f64(g.h)
        self._call("rsdft_gpu_init", g.N[0], g.N[1], g.N[2], _p(h), g.dV, system.Md, _p(coef),
                   system.coef_lap0)
        self.set_vloc(system.vloc if vloc is None else vloc)
        self.set_projectors(system.proj if proj is None else proj, with_derivatives)
        self._call("rsdft_gpu_set_solver_params", ncg, mloop)
        self.set_precision(precision)
        self._attached = None

    # -- plumbing ----------------------------------------------------------------
    def _call(self, name, *args):
        r = getattr(self.lib, name)(*args)
        if r != 0:
            msg = self.lib.rsdft_gpu_last_error().decode(errors="replace")
            raise RsdftGpuError(f"{name} -> {r}: {msg}")
        return r

    def set_precision(self, mode: int):
        self._call("rsdft_gpu_set_precision", int(mode))
        self.precision = int(mode)

    def set_precond_mode(self, mode: int):
        """1: s-step preconditioner (default), 0: step-by-step recurrence."""
        self._call("rsdft_gpu_set_precond_mode", int(mode))

    def set_eigensolver(self, mode: int):
        """1: refinement when Hsub is nearly diagonal, else mode 3 (default);
        3: LDS-resident tridiagonal reduction + bisection / inverse iteration;
        2: the reduction + dstedc; 0: always dsyevd."""
        self._call("rsdft_gpu_set_eigensolver", int(mode))

    def set_sweep_mode(self, mode: int):
        """1: fused Gram-Schmidt + Rayleigh-Ritz from CG's H x (default),
        0: gram_schmidt then subspace_diag as two steps."""
        self._call("rsdft_gpu_set_sweep_mode", int(mode))

    def set_vloc(self, vloc):
        v = _f64(vloc)
        self._call("rsdft_gpu_set_vloc", _p(v), 1)

    def set_projectors(self, pr, with_derivatives=True):
        self._tabs = (_i32(pr.MJJ), _i32(pr.JJP), _f64(pr.uVk), _i32(pr.iuV))
        MJJ, JJP, uVk, iuV = self._tabs
        self._call("rsdft_gpu_set_projectors", pr.nzlma, pr.MAXMJJ, _p(MJJ), _p(JJP), _p(uVk),
                   _p(iuV))
        if pr.nzlma == 0:            # no nonlocal part (and no forces)
            self.natom = 0
            return
        if with_derivatives:
            self._dtabs = (_i32(pr.amap), _i32(pr.MJJ_MAP), _i32(pr.JJ_MAP), _f64(pr.duVdR))
            amap, mjm, jjm, dud = self._dtabs
            natom = int(np.max(pr.amap))
            self._call("rsdft_gpu_set_projector_derivatives", natom, _p(amap), pr.MMJJ, _p(mjm),
                       _p(jjm), _p(dud))
        self.natom = int(np.max(pr.amap))

    # -- operators (hamiltonian_module.f90:20-72) --------------------------------
    def hamiltonian(self, tpsi):
        t = _f64(np.atleast_2d(tpsi))
        h = np.empty_like(t)
        self._call("rsdft_gpu_hamiltonian", 1, 1, _p(t), _p(h), 1, self.ML, 1, t.shape[0])
        return h

    def hamiltonian_noalloc(self, d_tpsi: int, d_htpsi: int, nb: int):
        """The allocation-free hamiltonian on DEVICE arrays (addresses of nb
        band-major FP64 bands each), FP64 arithmetic."""
        self._call("rsdft_gpu_hamiltonian_noalloc", 1, 1, ctypes.c_void_p(d_tpsi),
                   ctypes.c_void_p(d_htpsi), 1, self.ML, 1, int(nb))

    def op_kinetic(self, tpsi):
        t = _f64(np.atleast_2d(tpsi))
        h = np.empty_like(t)
        self._call("rsdft_gpu_op_kinetic", 1, _p(t), _p(h), 1, self.ML, 1, t.shape[0])
        return h

    def op_localpot(self, f, vf=None):
        f = _f64(np.atleast_2d(f))
        vf = np.zeros_like(f) if vf is None else _f64(vf).copy()
        self._call("rsdft_gpu_op_localpot", 1, self.ML, f.shape[0], _p(f), _p(vf))
        return vf

    def op_nonlocal(self, tpsi, htpsi=None):
        t = _f64(np.atleast_2d(tpsi))
        h = np.zeros_like(t) if htpsi is None else _f64(htpsi).copy()
        self._call("rsdft_gpu_op_nonlocal", 1, 1, _p(t), _p(h), 1, self.ML, 1, t.shape[0])
        return h

    # -- residency -----------------------------------------------------------------
    def attach(self, unk: np.ndarray):
        """Device mirror of unk (float64, C-contiguous (MB, ML)); kept by reference."""
        if unk.dtype != np.float64 or not unk.flags.c_contiguous:
            raise ValueError("attach needs a C-contiguous float64 array")
        self._call("rsdft_gpu_wf_attach", _p(unk), self.ML, unk.shape[0])
        self._attached = unk

    def sync(self, to_host: bool = True):
        self._call("rsdft_gpu_wf_sync", 0 if to_host else 1)

    def detach(self):
        self._call("rsdft_gpu_wf_detach")
        self._attached = None

    def _wf(self, unk):
        if self._attached is not None and unk is self._attached:
            return unk
        return _f64(unk).copy()

    # -- eigensolver -----------------------------------------------------------------
    def conjugate_gradient(self, unk):
        x = self._wf(unk)
        MB = x.shape[0]
        esp, res = np.zeros(MB), np.zeros(MB)
        self._call("rsdft_gpu_conjugate_gradient", 1, self.ML, MB, 1, 1, _p(x), _p(esp), _p(res))
        return x, esp, res

    def gram_schmidt(self, unk):
        x = self._wf(unk)
        self._call("rsdft_gpu_gram_schmidt", _p(x), self.ML, x.shape[0], 1, x.shape[0])
        return x

    def subspace_diag(self, unk):
        x = self._wf(unk)
        esp = np.zeros(x.shape[0])
        self._call("rsdft_gpu_subspace_diag", _p(x), _p(esp), self.ML, x.shape[0])
        return x, esp

    def sweep(self, unk, nsweep: int = 1):
        x = self._wf(unk)
        esp, res = np.zeros(x.shape[0]), np.zeros(x.shape[0])
        self._call("rsdft_gpu_sweep", _p(x), _p(esp), _p(res), self.ML, x.shape[0], nsweep)
        return x, esp, res

    # -- adjoints ----------------------------------------------------------------------
    def calc_density(self, unk, occ):
        x = self._wf(unk)
        occ = _f64(occ)
        rho = np.zeros(self.ML)
        self._call("rsdft_gpu_calc_density", _p(x), self.ML, x.shape[0], _p(occ), _p(rho))
        return rho

    # -- SCF pieces (scf_module.f90:146-366) ---------------------------------------
    def calc_hartree(self, rho):
        """(Vh, E_H) of the density rho (hartree_sol_module.f90:19-122)."""
        r = _f64(rho)
        vh = np.zeros(self.ML)
        eh = np.zeros(1)
        self._call("rsdft_gpu_calc_hartree", _p(r), _p(vh), _p(eh))
        return vh, float(eh[0])

    def calc_xc(self, rho):
        """(Vxc, Exc, Ex, Ec) of the PZ81 LDA (xc_ldapz81_module.f90:17-75)."""
        r = _f64(rho)
        v = np.zeros(self.ML)
        e = np.zeros(3)
        self._call("rsdft_gpu_calc_xc", _p(r), _p(v), _p(e))
        return v, float(e[0]), float(e[1]), float(e[2])

    def scf_init(self, vion, rho, nelectron, mmix=4, beta=1.0, ekbt=1e-5, kinteg=5):
        """init_scf + init_mixing for imix = 10; Vloc = Vion + Vh[rho] + Vxc[rho]."""
        self._scf_in = (_f64(vion), _f64(rho))
        self._call("rsdft_gpu_scf_init", _p(self._scf_in[0]), _p(self._scf_in[1]),
                   float(nelectron), int(mmix), float(beta), float(ekbt), int(kinteg))

    def scf_iterate(self, unk):
        """One scf iteration: returns (psi, esp, occ, res, info dict)."""
        x = self._wf(unk)
        MB = x.shape[0]
        esp, occ, res, info = np.zeros(MB), np.zeros(MB), np.zeros(MB), np.zeros(10)
        self._call("rsdft_gpu_scf_iterate", _p(x), self.ML, MB, _p(esp), _p(occ), _p(res),
                   _p(info))
        keys = ("etot", "ehwf", "eeig", "eloc", "eion", "eh", "exc", "efermi", "sqerr_rho",
                "sqerr_v")
        return x, esp, occ, res, dict(zip(keys, (float(v) for v in info)))

    def scf_fields(self):
        """(rho, Vloc, Vh, Vxc) after the last SCF step."""
        f = [np.zeros(self.ML) for _ in range(4)]
        self._call("rsdft_gpu_scf_get_fields", *(_p(a) for a in f))
        return tuple(f)

    # -- complex wavefunctions at one k point --------------------------------------
    def set_kpoint(self, kvec, proj):
        """The k point's kinetic terms and Bloch-phased projector tables (complex
        ``proj.uVk``), e.g. from ``system.make_projectors(..., kvec=kvec)``."""
        from .system import kpoint_kinetic
        g = self.sys.grid
        ck2, nabk = kpoint_kinetic(self.sys.Md, g.h, kvec)
        self._kp_nabk = _f64(np.asarray(nabk).T.ravel())          # Fortran (3, Md)
        u = np.ascontiguousarray(np.asarray(proj.uVk, dtype=np.complex128))
        self._kp_tabs = (_i32(proj.MJJ), _i32(proj.JJP), u.view(np.float64), _i32(proj.iuV))
        MJJ, JJP, uv, iuV = self._kp_tabs
        self._call("rsdft_gpu_set_kpoint", float(ck2), _p(self._kp_nabk), proj.nzlma,
                   proj.MAXMJJ, _p(MJJ), _p(JJP), _p(uv), _p(iuV))

    @staticmethod
    def _z(a):
        return np.ascontiguousarray(np.atleast_2d(np.asarray(a, dtype=np.complex128))).copy()

    def zhamiltonian(self, psi):
        t = self._z(psi)
        h = np.empty_like(t)
        self._call("rsdft_gpu_zhamiltonian", 1, 1, _p(t.view(np.float64)), _p(h.view(np.float64)),
                   1, self.ML, 1, t.shape[0])
        return h

    def zconjugate_gradient(self, psi):
        x = self._z(psi)
        esp, res = np.zeros(x.shape[0]), np.zeros(x.shape[0])
        self._call("rsdft_gpu_zconjugate_gradient", _p(x.view(np.float64)), _p(esp), _p(res),
                   self.ML, x.shape[0])
        return x, esp, res

    def zgram_schmidt(self, psi):
        x = self._z(psi)
        self._call("rsdft_gpu_zgram_schmidt", _p(x.view(np.float64)), self.ML, x.shape[0])
        return x

    def zsubspace_diag(self, psi):
        x = self._z(psi)
        esp = np.zeros(x.shape[0])
        self._call("rsdft_gpu_zsubspace_diag", _p(x.view(np.float64)), _p(esp), self.ML,
                   x.shape[0])
        return x, esp

    def zsweep(self, psi, nsweep: int = 1):
        x = self._z(psi)
        esp, res = np.zeros(x.shape[0]), np.zeros(x.shape[0])
        self._call("rsdft_gpu_zsweep", _p(x.view(np.float64)), _p(esp), _p(res), self.ML,
                   x.shape[0], int(nsweep))
        return x, esp, res

    def calc_force_ps_nloc2(self, unk, occ):
        x = self._wf(unk)
        occ = _f64(occ)
        f = np.zeros(3 * self.natom)
        self._call("rsdft_gpu_force_ps_nloc2", _p(x), self.ML, x.shape[0], _p(occ), _p(f))
        return f.reshape(self.natom, 3)

    def band_energy_grad(self, unk, occ):
        """(E, dE/dVloc (ML,), dE/dR (natom, 3)) at fixed psi."""
        x = self._wf(unk)
        occ = _f64(occ)
        E = np.zeros(1)
        dV = np.zeros(self.ML)
        dR = np.zeros(3 * self.natom)
        self._call("rsdft_gpu_band_energy_grad", _p(x), self.ML, x.shape[0], _p(occ), _p(E),
                   _p(dV), _p(dR))
        return float(E[0]), dV, dR.reshape(self.natom, 3)

    # -- measurement ---------------------------------------------------------------------
    def test_eigh(self, H, mode: int = 2, vectors: bool = True):
        """(w, V, ms): eigenpairs of the symmetric matrix H by eigensolver `mode`."""
        A = np.asfortranarray(_f64(H))
        n = A.shape[0]
        w = np.zeros(n)
        V = np.zeros((n, n), order="F") if vectors else None
        ms = np.zeros(1)
        self._call("rsdft_gpu_test_eigh", _p(A), n, int(mode), _p(w),
                   _p(V) if vectors else None, _p(ms))
        return w, V, float(ms[0])

    def test_hpsi(self, nrep: int = 10):
        """ms per all-band application: hamiltonian, kinetic, localpot, nonlocal,
        stencil kernel of the hamiltonian, projection kernel, nonlocal term
        stage, plain copy of all bands."""
        t = np.zeros(8)
        self._call("rsdft_gpu_test_hpsi", nrep, _p(t))
        return t

    def test_gemm(self, nrep: int = 5):
        """ms per call of the sweep's tall GEMMs on the attached psi:
        X^T X, X^T Y (one triangle), X U."""
        t = np.zeros(3)
        self._call("rsdft_gpu_test_gemm", nrep, _p(t))
        return t

    def timers(self):
        t = np.zeros(10)
        self._call("rsdft_gpu_get_timers", _p(t), 10)
        return {"hamiltonian": t[0], "precond": t[1], "cg": t[2], "gram_schmidt": t[3],
                "subspace_diag": t[4], "sweep": t[5], "hpsi_bands": t[6], "eig_refined": t[7],
                "cg_rejected_bands": t[8], "pc_recurrence_bands": t[9]}

    def reset_timers(self):
        self._call("rsdft_gpu_reset_timers")

    def finalize(self):
        self._call("rsdft_gpu_finalize")
