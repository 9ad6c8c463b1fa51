"""The SCF iteration on the GPU (scf_module.f90:146-366, one spin, Gamma
point) against the oracle's restatement (oracle/scf.py), which tests/
test_oracle_scf.py pins by closed forms:

* calc_hartree (hartree_sol_module.f90:19-122): the dense-DFT Poisson solve
  against numpy's FFT on the G sphere, and a plane wave inside the sphere
  against 4 pi / G^2 directly; 1e-12 relative;
* calc_xc, PZ81 LDA (xc_ldapz81_module.f90:17-210): 1e-13 relative;
* calc_fermi, Methfessel-Paxton (fermi_mp_module.f90:96-287): occupations
  and Fermi level to 1e-12;
* six SCF iterations of Si8 (32 bands, 64 electrons) from the density of
  the orthonormalised seeded start: eigenvalues, total and Harris energies,
  the two mixing errors and the final density, with the Pulay step active
  from the fourth iteration (mmix = 4); eigenvalues and energies to 1e-9
  Ha, density 1e-10 relative.
"""
import math

import numpy as np
import pytest

from rsdft_amd.gpu import RsdftGpu, PREC_FP64, calc_fermi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(si8, gpu_ctx):
    d = RsdftGpu(si8, precision=PREC_FP64)
    yield d
    d.finalize()


@pytest.fixture(scope="module")
def start(si8, oracle_si8):
    x = oracle_si8.gram_schmidt(si8.psi0())
    return x, oracle_si8.density(x, si8.occ)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_hartree_matches_fft(dev, si8, start):
    from oracle import scf
    print("case: hartree si8 density of the seeded start", flush=True)
    _, rho = start
    vh, eh = dev.calc_hartree(rho)
    vo, eo = scf.hartree(rho, si8.grid)
    print(f"  rel err {rel(vh, vo):.2e}, E_H {eh:.12f} vs {eo:.12f}", flush=True)
    assert rel(vh, vo) < 1e-12
    assert abs(eh - eo) < 1e-12 * abs(eo)


@pytest.mark.parametrize("n", [(1, 0, 0), (2, -3, 1), (5, 7, -6)])
def test_hartree_plane_wave(dev, si8, n):
    print(f"case: hartree plane wave {n}", flush=True)
    g = si8.grid
    x = [np.arange(N) * (L / N) for N, L in zip(g.N, g.L)]
    X3, X2, X1 = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    G = [2 * math.pi * n[d] / g.L[d] for d in range(3)]
    rho = np.cos(G[0] * X1 + G[1] * X2 + G[2] * X3).ravel()
    g2 = sum(t * t for t in G)
    vh, eh = dev.calc_hartree(rho)
    assert np.max(np.abs(vh - 4 * math.pi / g2 * rho)) < 1e-12 * 4 * math.pi / g2
    omega = g.ML * g.dV
    assert abs(eh - math.pi * omega / g2) < 1e-10 * abs(eh)


def test_xc_matches_oracle(dev, si8, start):
    from oracle import scf
    print("case: PZ81 LDA si8", flush=True)
    _, rho = start
    # both branches of the correlation: rs < 1 at the densest points, >= 1 elsewhere
    rho = rho * np.linspace(0.2, 30.0, rho.size)
    v, exc, ex, ec = dev.calc_xc(rho)
    vo, exco, exo, eco = scf.xc_pz81(rho, si8.grid.dV)
    print(f"  rel err {rel(v, vo):.2e}, Exc {exc:.12f} vs {exco:.12f}", flush=True)
    assert rel(v, vo) < 1e-13
    for a, b in ((exc, exco), (ex, exo), (ec, eco)):
        assert abs(a - b) < 1e-12 * abs(b)


@pytest.mark.parametrize("kinteg", [0, 1, 5])
def test_fermi_matches_oracle(kinteg):
    from oracle import scf
    rng = np.random.default_rng(kinteg)
    esp = np.sort(rng.uniform(-0.3, 0.6, 40))
    for nel, ekbt in ((31.0, 1e-3), (64.0, 1e-5), (50.0, 2e-2)):
        occ, ef = calc_fermi(esp, nel, ekbt, kinteg)
        oo, eo = scf.calc_fermi(esp, nel, ekbt, kinteg)
        assert abs(ef - eo) < 1e-12
        assert np.max(np.abs(occ - oo)) < 1e-12
        assert abs(occ.sum() - nel) < 1e-8


def test_scf_iterations_match_oracle(dev, si8, start):
    from oracle import scf
    x0, rho0 = start
    nel = float(np.sum(si8.occ))
    run = scf.Scf(si8, rho0, nelectron=nel)
    dev.scf_init(si8.vloc, rho0, nel)
    # init_scf: the input density and the potentials of it
    rho, vloc, vh, vxc = dev.scf_fields()
    print(f"case: scf init si8: rho {rel(rho, rho0):.1e}, Vloc {rel(vloc, run.vloc):.1e}, "
          f"Vh {rel(vh, run.vh):.1e}, Vxc {rel(vxc, run.vxc):.1e}", flush=True)
    assert rel(rho, rho0) == 0.0
    for a, b in ((vloc, run.vloc), (vh, run.vh), (vxc, run.vxc)):
        assert rel(a, b) < 1e-12
    from oracle.oracle import Oracle
    h = dev.hamiltonian(x0[:4])
    ho = Oracle(si8, vloc=run.vloc).hamiltonian(x0[:4])
    print(f"  hamiltonian with the SCF potential: rel err {rel(h, ho):.1e}", flush=True)
    assert rel(h, ho) < 1e-12
    xg = x0.copy()
    xo = x0.copy()
    for it in range(6):
        print(f"case: scf iteration {it + 1} si8", flush=True)
        xg, eg, og, _, ig = dev.scf_iterate(xg)
        xo, eo, oo, io = run.iterate(xo)
        de = float(np.max(np.abs(eg - eo)))
        if de > 1e-9:
            print(f"  eig gpu {eg[:6]}\n  eig ref {eo[:6]}\n  occ gpu {og[:4]} ref {oo[:4]}\n"
                  f"  {ig}\n  {io}", flush=True)
        print(f"  max |d eig| {de:.2e} Ha, Etot {ig['etot']:.10f} vs {io['etot']:.10f}, "
              f"Ehwf {ig['ehwf']:.10f} vs {io['ehwf']:.10f}, sqerr_rho {ig['sqerr_rho']:.3e}",
              flush=True)
        assert de < 1e-9
        assert np.max(np.abs(og - oo)) < 1e-9
        for k in ("etot", "ehwf", "eeig", "eloc", "eion", "eh", "exc"):
            assert abs(ig[k] - io[k]) < 1e-9, k
        for k in ("sqerr_rho", "sqerr_v"):
            assert abs(ig[k] - io[k]) <= 1e-8 * abs(io[k]) + 1e-20, k
    rho, vloc, vh, vxc = dev.scf_fields()
    print(f"  final density rel err {rel(rho, run.rho):.2e}", flush=True)
    assert rel(rho, run.rho) < 1e-10
    assert rel(vloc, run.vloc) < 1e-10
    assert abs(np.sum(rho) * si8.grid.dV - nel) < 1e-10 * nel
    # the hamiltonian now runs with the new potential
    h = dev.hamiltonian(xg[:4])
    ho = Oracle(si8, vloc=run.vloc).hamiltonian(xg[:4])
    assert rel(h, ho) < 1e-11
