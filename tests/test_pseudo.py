"""Host set-up of the real pseudopotential path (rsdft_amd/pseudo.py) against
what the reference printed for its shipped Si216 example (fixtures in
tests/golden/si216_psv/, made by tests/golden/make_si216_psv_fixture.py from
examples/Si216/a.log, fort.99, fort.970 and Si_psv.dat) and against closed
forms:

* read_PSV's summary line sum(cdd) = 3.99144241688915, anorm, Mr, the file's
  erf split (ps_read_psv.f90:340-356, a.log:95-111);
* the G sphere: Ecut, Gcut, 54,435 vectors in 462 shells (a.log:171-174);
* the Ewald energy -907.161779580626444 (a.log:302), independent of the split
  parameter eta;
* the initial density's sum / min / max before normalisation (a.log:283);
* the filtering kernel int_0^qc j_L(qr) j_L(qr') q^2 dq in its closed forms
  (filtering.f90) against quadrature of scipy's spherical Bessel functions
  for L = 0..4, on every branch (r = 0, r' = 0, r' = r);
* the mask function's defining eigen-equation (maskf_module.f90);
* Simpson's rule and Neville interpolation on polynomials they integrate /
  reproduce exactly; the real Y_lm against the oracle's restatement of
  ylm_module.f90 and their orthonormality;
* the local form factor's independence of the erf split.
The whole chain (projector tables, Vion, the SCF on the GPU) is pinned by the
reference's final total energy and eigenvalues in tests/test_gpu_psv_si216.py.
"""
import json
import math
import os

import numpy as np
import pytest

from rsdft_amd import pseudo
from rsdft_amd.system import Grid

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si216_psv")


@pytest.fixture(scope="module")
def ref():
    return json.load(open(os.path.join(GOLD, "results.json")))


@pytest.fixture(scope="module")
def atoms():
    return json.load(open(os.path.join(GOLD, "atoms.json")))


@pytest.fixture(scope="module")
def ps():
    return pseudo.read_psv(os.path.join(GOLD, "Si_psv.dat.gz"))


@pytest.fixture(scope="module")
def grid(atoms):
    return Grid(tuple(atoms["Ngrid"]), (atoms["AX"],) * 3)


def test_read_psv_summary(ps, ref):
    assert ps.mr == ref["psv_mr"]
    s = float(np.sum(ps.cdd * ps.rab))
    assert abs(s - ref["psv_sum_cdd"]) < 1e-13 * ref["psv_sum_cdd"]
    assert np.allclose(ps.anorm, ref["psv_anorm"], rtol=0, atol=1e-15)
    assert np.allclose(ps.parloc, ref["psv_parloc"], rtol=0, atol=1e-15)
    assert ps.zps == 4.0 and ps.zatom == 14.0 and ps.xc == "LDAPZ81"
    assert list(ps.lo) == [0, 1] and list(ps.inorm) == [1.0, 1.0]
    assert ps.rad[0] == 0.0 and ps.viod[0, 0] == 0.0
    # rab is r times the log-mesh step: the mesh integral of a Gaussian
    g = np.exp(-ps.rad ** 2)
    assert abs(pseudo.simpson(ps.rad ** 2 * g * ps.rab) - math.sqrt(math.pi) / 4) < 1e-10


def test_g_sphere(grid, ref):
    gs = pseudo.g_sphere(grid)
    assert len(gs.g2) == ref["ngsphere"]
    assert len(gs.gg) == ref["nmgl"]
    assert abs(gs.ecut - ref["ecut"]) < 1e-13 * ref["ecut"]
    assert abs(gs.gcut - ref["gcut"]) < 1e-13 * ref["gcut"]
    assert gs.gg[0] == 0.0 and np.all(np.diff(gs.gg) > 1e-12)
    assert np.allclose(gs.gg[gs.shell], gs.g2, rtol=1e-12, atol=0)


def test_ewald_matches_reference_and_is_eta_independent(atoms, ref):
    cell = np.eye(3) * atoms["AX"]
    frac = np.array(atoms["frac"])
    z = np.full(len(frac), 4.0)
    e = pseudo.ewald_energy(cell, frac, z)
    print(f"Ewald {e:.15f} vs {ref['eewald']:.15f}")
    assert abs(e - ref["eewald"]) < 1e-12 * abs(ref["eewald"])
    for eta in (0.004, 0.02):
        assert abs(pseudo.ewald_energy(cell, frac, z, eta=eta) - e) < 1e-12 * abs(e)


def test_ewald_closed_form_simple_cubic():
    """One unit charge per cubic cell: the Madelung energy -1.4186487/L x Z^2/2."""
    L = 7.0
    e = pseudo.ewald_energy(np.eye(3) * L, np.zeros((1, 3)), np.array([1.0]))
    assert abs(e - (-2.837297479480620 / (2 * L))) < 1e-12


def test_initial_density_matches_reference(ps, grid, atoms, ref):
    gs = pseudo.g_sphere(grid)
    frac = np.array(atoms["frac"])
    vcell = grid.ML * grid.dV
    cddg = pseudo.initrho_form_factor(ps, gs.gg, vcell)
    rho = pseudo.to_real_space(grid, gs, cddg[gs.shell] * pseudo.structure_factor(gs, frac))
    s, lo, hi = float(np.sum(rho) * grid.dV), float(rho.min()), float(rho.max())
    rs, rlo, rhi = ref["initrho_before"]
    print(f"initrho sum {s:.10f} min {lo:.10f} max {hi:.10f} vs {rs} {rlo} {rhi}")
    assert abs(s - rs) < 1e-9 and abs(lo - rlo) < 1e-10 and abs(hi - rhi) < 1e-10
    # the G = 0 term carries the whole charge of the form factor
    assert abs(cddg[0] * vcell * len(frac) - s) < 1e-9 * s


@pytest.mark.parametrize("L", [0, 1, 2, 3, 4])
def test_filter_kernel_closed_forms(L):
    from scipy.integrate import quad
    from scipy.special import spherical_jn
    qc = 4.8
    rng = np.random.default_rng(L)
    cases = [(r, rng.uniform(0.05, 4.0, 4)) for r in rng.uniform(0.05, 4.0, 3)]
    r_eq = 1.37
    cases.append((r_eq, np.array([r_eq, 0.0, 2.2])))
    cases.append((0.0, np.array([0.0, 0.3, 1.9])))
    for r, r1 in cases:
        t = pseudo._filter_kernel(L, qc, r, r1)
        for k, y in enumerate(r1):
            want = quad(lambda q: spherical_jn(L, q * r) * spherical_jn(L, q * y) * q * q,
                        0.0, qc, epsabs=1e-13, epsrel=1e-12, limit=200)[0]
            got = 0.0 if t is None else t[k]
            assert abs(got - want) < 1e-9 * max(1.0, abs(want)), (L, r, y, got, want)


def test_mask_function_eigen_equation():
    eta = 8.0
    xm, mk = pseudo.mask_function(eta)
    n = pseudo.NMSK - 1
    dx = 1.0 / n
    x = np.arange(1, n + 1) * dx
    v = mk[1:n + 1] * x                          # the eigenvector up to scale
    s, d = x[:, None] + x[None, :], x[:, None] - x[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        K = np.sin(s * eta) / s - np.where(d == 0, 0.0, np.sin(d * eta) / np.where(d == 0, 1, d))
    K[np.diag_indices(n)] = np.sin(2 * x * eta) / (2 * x) + n * math.pi - eta
    lam = v @ K @ v / (v @ v)
    assert np.linalg.norm(K @ v - lam * v) < 1e-10 * np.linalg.norm(K) * np.linalg.norm(v)
    assert abs(lam - np.linalg.eigvalsh(K)[0]) < 1e-9 * abs(lam)
    assert mk[0] == 1.0 and mk[-1] == 0.0
    assert abs(mk[1] - 1.0) < 1e-2 and np.all(np.abs(mk[1:n]) < 1.2)
    vals = pseudo.mask_at(np.array([0.0, 0.004, xm[7], 0.5]), xm, mk)
    assert vals[0] == 1.0 and vals[1] == 1.0 and abs(vals[2] - mk[7]) < 1e-14


def test_simpson_and_neville_exactness():
    x = np.linspace(0.0, 2.0, 41)                 # odd count: Simpson exact for cubics
    f = 3 * x ** 3 - x ** 2 + 2
    h = x[1] - x[0]
    assert abs(pseudo.simpson(f) * h - (3 * 4 - 8 / 3 + 4)) < 1e-12
    # even count: the last point is left out (panels up to n-1)
    assert abs(pseudo.simpson(np.append(f, 1e6)) - pseudo.simpson(f)) == 0.0
    xa = np.sort(np.random.default_rng(1).uniform(0, 1, (5, 6)), axis=1)
    ya = 1 - 2 * xa + 0.5 * xa ** 3 - xa ** 5
    xq = np.full(5, 0.37)
    v, e = pseudo.neville(xa, ya, xq)
    assert np.max(np.abs(v - (1 - 2 * 0.37 + 0.5 * 0.37 ** 3 - 0.37 ** 5))) < 1e-13


def test_real_ylm_matches_reference_restatement_and_is_orthonormal():
    from oracle.ylm import ylm
    rng = np.random.default_rng(3)
    p = rng.normal(size=(50, 3))
    p[0] = (0.0, 0.7, -0.2)                       # x = 0: the reference's azimuth branch
    p[1] = (-0.3, 0.0, 0.0)
    for L in range(5):
        for m in range(-L, L + 1):
            got = pseudo.real_ylm(p[:, 0], p[:, 1], p[:, 2], L, m)
            want = np.array([ylm(*q, L, m) for q in p])
            assert np.max(np.abs(got - want)) < 1e-13, (L, m)
    # orthonormality on a Gauss-Legendre x uniform-azimuth product rule
    ct, w = np.polynomial.legendre.leggauss(12)
    ph = np.arange(24) * 2 * math.pi / 24
    C, P = np.meshgrid(ct, ph, indexing="ij")
    W = np.repeat(w[:, None], 24, axis=1) * (2 * math.pi / 24)
    st = np.sqrt(1 - C ** 2)
    X, Y, Z = st * np.cos(P), st * np.sin(P), C
    Ys = [pseudo.real_ylm(X, Y, Z, L, m) for L in range(4) for m in range(-L, L + 1)]
    M = np.array([[np.sum(a * b * W) for b in Ys] for a in Ys])
    assert np.max(np.abs(M - np.eye(len(Ys)))) < 1e-12
    # the closed form of the p channels (Condon-Shortley sign)
    q = p[2:]
    r = np.linalg.norm(q, axis=1)
    k1 = math.sqrt(3 / (4 * math.pi))
    assert np.allclose(pseudo.real_ylm(*q.T, 1, 1), -k1 * q[:, 0] / r, atol=1e-14)
    assert np.allclose(pseudo.real_ylm(*q.T, 1, -1), -k1 * q[:, 1] / r, atol=1e-14)
    assert np.allclose(pseudo.real_ylm(*q.T, 1, 0), k1 * q[:, 2] / r, atol=1e-14)


def test_local_form_factor_split_independent(ps, grid):
    gs = pseudo.g_sphere(grid)
    vcell = grid.ML * grid.dV
    a = pseudo.local_form_factor(ps, gs.gg, vcell)
    b = pseudo.local_form_factor(ps, gs.gg, vcell, parloc=np.array([0.7, 2.0, 0.3, 0.4]))
    assert np.max(np.abs(a - b)) < 1e-11 * np.max(np.abs(a))
    # the G -> 0 limit is finite and the form factor tends to -4 pi Z / (V G^2)
    big = gs.gg > 15.0
    assert np.all(np.isfinite(a))
    tail = a[big] * gs.gg[big] * vcell / (-4 * math.pi * ps.zps)
    assert np.all(np.abs(tail) < 1.5)


def test_filtered_projectors_and_tables(ps, grid, atoms):
    gs = pseudo.g_sphere(grid)
    fp = pseudo.filter_projectors(ps, gs.gcut)
    # the cut-off grows to the first log-mesh point at or beyond 1.5 Rc
    k = int(np.argmin(np.abs(ps.rad - 1.5 * ps.rps[0])))
    k = k + 1 if ps.rad[k] < 1.5 * ps.rps[0] else k
    assert fp.rps[0] == ps.rad[k]
    assert (fp.nrps[0] - 2) * pseudo.DR_FILTER < fp.rps[0] <= (fp.nrps[0] - 1) * pseudo.DR_FILTER
    # the p projector vanishes at r = 0 and both end at the (masked) cut-off
    assert fp.viod[0, 1] == 0.0
    assert abs(fp.viod[fp.nrps[0] - 1, 0]) < 1e-3 * np.max(np.abs(fp.viod[:, 0]))
    # the file holds r beta(r) on the log mesh, the filtered table beta(r);
    # filtering keeps the norm int beta^2 r^2 dr within 1 %
    n_orig = pseudo.simpson(ps.viod[:, 0] ** 2 * ps.rab)
    r1 = fp.rad1[:fp.nrps[0]]
    n_filt = pseudo.simpson((fp.viod[:fp.nrps[0], 0] * r1) ** 2) * pseudo.DR_FILTER
    assert abs(n_filt / n_orig - 1.0) < 1e-2
    frac = np.array(atoms["frac"])
    pr = pseudo.projector_tables(grid, frac, fp, np.eye(3) * atoms["AX"])
    assert pr.nzlma == 4 * len(frac)
    assert list(pr.lmap[:4]) == [0, 1, 1, 1] and list(pr.mmap[:4]) == [0, -1, 0, 1]
    # points sorted, unique, inside the grid; the s projector of atom 0 is
    # spherically symmetric: its grid norm matches the radial one to 2 %
    for z in range(8):
        j = pr.JJP[z, :pr.MJJ[z]]
        assert np.all(np.diff(j) > 0) and j[0] >= 1 and j[-1] <= grid.ML
    s0 = np.sum(pr.uVk[0, :pr.MJJ[0]] ** 2) * grid.dV
    assert abs(s0 / n_filt - 1.0) < 2e-2
