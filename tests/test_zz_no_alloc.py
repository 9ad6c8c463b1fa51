"""The allocation-free hamiltonian (rsdft_gpu_hamiltonian_noalloc, device
arrays, projection coefficients in the library's static device scratch)
against the allocating entry point rsdft_gpu_hamiltonian in FP64 on the same
bands: two calls in a row, the second with fewer bands (the scratch keeps the
first call's contents), then the size check.  Device memory comes from the
HIP runtime the library itself links, looked up through the library's own
handle (its dependencies' symbols), so no second runtime is loaded even when
another one (torch's) is already in the process."""
import ctypes

import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, RsdftGpuError, PREC_FP64

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


class _Runtime:
    """hipMalloc / hipFree / hipMemcpy / hipDeviceSynchronize of the library's runtime."""

    def __init__(self, lib):
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            setattr(self, name, f)


def hip_runtime():
    from rsdft_amd.gpu import load_library
    hip = _Runtime(load_library())
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def dmalloc(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def test_noalloc_hamiltonian_matches_allocating_entry(si8, gpu_ctx):
    hip = hip_runtime()
    dev = RsdftGpu(si8, precision=PREC_FP64)
    ML = si8.grid.ML
    nmax = 12
    d_in = dmalloc(hip, nmax * ML * 8)
    d_out = dmalloc(hip, nmax * ML * 8)
    try:
        for b0, nb in ((3, 12), (20, 5)):
            print(f"case: noalloc hamiltonian si8 bands {b0}..{b0 + nb - 1}", flush=True)
            psi = np.ascontiguousarray(si8.psi0(b0, nb))
            ref = dev.hamiltonian(psi)
            got = np.full_like(psi, np.nan)
            assert hip.hipMemcpy(d_in, psi.ctypes.data, psi.nbytes, HIP_H2D) == 0
            assert hip.hipMemcpy(d_out, got.ctypes.data, got.nbytes, HIP_H2D) == 0
            dev.hamiltonian_noalloc(d_in.value, d_out.value, nb)
            assert hip.hipMemcpy(got.ctypes.data, d_out, got.nbytes, HIP_D2H) == 0
            err = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
            print(f"  rel diff {err:.3e} (bit-identical: {np.array_equal(got, ref)})", flush=True)
            assert np.all(np.isfinite(got))
            assert err <= 1e-14
        # beyond the scratch: refused on the host before any launch
        with pytest.raises(RsdftGpuError, match="exceed"):
            dev.hamiltonian_noalloc(d_in.value, d_out.value, (2 << 20) // si8.proj.nzlma + 1)
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(d_in)
        hip.hipFree(d_out)
        dev.finalize()
