"""Stream-unformatted input/output files of the Fortran sweep driver
(rsdft_amd/fortran/sweep_driver.f90), in the reference's array layouts."""
from __future__ import annotations

import numpy as np


def write_driver_input(path, system, psi, nsweep):
    g = system.grid
    pr = system.proj
    with open(path, "wb") as f:
        np.array(list(g.N) + [system.Md, pr.nzlma, pr.MAXMJJ, psi.shape[0], nsweep],
                 dtype=np.int32).tofile(f)
        np.array(list(g.h) + [g.dV, system.coef_lap0], dtype=np.float64).tofile(f)
        np.ascontiguousarray(np.asarray(system.coef_lap).T, dtype=np.float64).tofile(f)
        np.asarray(system.vloc, dtype=np.float64).tofile(f)
        np.asarray(pr.MJJ, dtype=np.int32).tofile(f)
        np.ascontiguousarray(pr.JJP, dtype=np.int32).tofile(f)
        np.ascontiguousarray(pr.uVk, dtype=np.float64).tofile(f)
        np.asarray(pr.iuV, dtype=np.int32).tofile(f)
        np.ascontiguousarray(psi, dtype=np.float64).tofile(f)


def read_driver_output(path, ML, MB):
    a = np.fromfile(path, dtype=np.float64)
    o = 0
    hpsi = a[o:o + ML * MB].reshape(MB, ML); o += ML * MB
    esp = a[o:o + MB]; o += MB
    res = a[o:o + MB]; o += MB
    psi = a[o:o + ML * MB].reshape(MB, ML); o += ML * MB
    hpsi_omp = a[o:o + ML * MB].reshape(MB, ML)
    return hpsi, esp, res, psi, hpsi_omp
