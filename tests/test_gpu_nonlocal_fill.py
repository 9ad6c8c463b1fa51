"""H psi with the nonlocal term written in grid order (k_nl_fill) and the
stencil accumulating onto it, against the FP64 oracle: clusters of atoms whose
projector spheres cover one grid point up to 4 times (the fill's largest
membership layout) and 6 times (beyond it: the scatter-add apply is used).
Bounds as in the Si8 hamiltonian tests: max |diff| / max |H psi| < 2e-6 (FP32
storage) and < 1e-12 (FP64).
"""
import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64, PREC_MIXED_FP32
from rsdft_amd.system import make_custom

pytestmark = pytest.mark.gpu

# atom offsets (bohr) around the cell centre: a tetrahedron, then two more
OFFSETS = 0.6 * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1],
                          [2, 0, 0], [0, 2, 0]], dtype=np.float64)


def max_memberships(system):
    """Largest number of atoms whose sphere (union over the atom's channels)
    holds one grid point."""
    p = system.proj
    cnt = np.zeros(system.grid.ML, np.int64)
    for a in range(1, system.natom + 1):
        sel = np.nonzero(p.amap == a)[0]
        pts = np.unique(np.concatenate([p.JJP[l, :p.MJJ[l]] for l in sel])) - 1
        cnt[pts] += 1
    return int(cnt.max())


@pytest.mark.parametrize("prec,name,tol", [(PREC_MIXED_FP32, "FP32", 2e-6),
                                           (PREC_FP64, "FP64", 1e-12)])
@pytest.mark.parametrize("natom", [4, 6])
def test_hamiltonian_overlapping_spheres(gpu_ctx, natom, prec, name, tol):
    from oracle.oracle import Oracle
    centre = np.full(3, 5.13)
    s = make_custom((24, 24, 24), (10.26,) * 3, centre + OFFSETS[:natom], nband=8)
    assert max_memberships(s) == natom
    print(f"case: {name} hamiltonian, {natom} overlapping spheres, 24^3", flush=True)
    dev = RsdftGpu(s, precision=prec)
    psi = s.psi0(0, 8)
    h = dev.hamiltonian(psi)
    ref = Oracle(s).hamiltonian(psi)
    err = float(np.max(np.abs(h - ref)) / np.max(np.abs(ref)))
    print(f"  rel err {err:.2e}", flush=True)
    assert err < tol, err
