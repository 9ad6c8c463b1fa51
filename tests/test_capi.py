"""CPU checks of the boundary: the C-ABI library loads without a GPU and
exports every entry point include/rsdft_gpu.h declares; the host-side input
generator keeps the reference's layout conventions."""
import os
import re

import numpy as np

from rsdft_amd import gpu
from rsdft_amd import system as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rsdft_gpu.h")).read()
    return sorted(set(re.findall(r"RSDFT_API\s+[\w\s\*]*?\b(rsdft_gpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = gpu.load_library()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(gpu.EXPORTS)


def test_library_reports_errors_without_init():
    lib = gpu.load_library()
    import ctypes
    buf = np.zeros(8)
    r = lib.rsdft_gpu_wf_sync(0)
    assert r != 0
    assert b"" != lib.rsdft_gpu_last_error()


def test_config_sizes_match_the_workloads():
    s = S.make_config("si8")
    assert s.grid.N == (24, 24, 24) and s.nband == 32 and s.natom == 8
    assert abs(s.grid.h[0] - 0.4275) < 1e-15
    assert s.proj.nzlma == 32 and int(s.occ.sum()) == 32
    s64 = S.make_config("si64", nband=4)
    assert s64.natom == 64 and s64.grid.N == (48, 48, 48)
    assert np.allclose(s64.positions[0], [0.1, 0.05, 0.0])


def test_psi0_is_reproducible_and_piecewise():
    s = S.make_config("si8")
    a = s.psi0()
    b = s.psi0(5, 3)
    assert np.array_equal(a[5:8], b)
    assert -1.0 <= a.min() and a.max() < 1.0
    assert abs(a.mean()) < 1e-2


def test_vloc_is_a_sum_of_twenty_plane_waves():
    s = S.make_config("si8")
    v, n, phi = S.make_vloc(s.grid, s.seed)
    assert n.shape == (20, 3) and np.all(np.sum(n * n, axis=1) <= 9) and np.all(np.any(n != 0, axis=1))
    f = np.fft.fftn(v.reshape(s.grid.N[::-1]))
    assert np.sum(np.abs(f) > 1e-8 * np.abs(f).max()) <= 40      # +/- G of 20 terms
    assert abs(np.mean(v)) < 1e-12
