"""The oracle's complex k-point restatement (oracle/kpoint.py) pinned without
itself:

* at k = 0 on real wavefunctions it reproduces the real oracle (a separate C
  restatement of the reference) operator by operator and through CG,
  Gram-Schmidt, subspace diagonalisation and sweeps;
* the kinetic operator with zcoef_kin has plane waves exp(iq.r) as
  eigenvectors with the closed-form eigenvalue
  c0 + |k|^2/2 + sum_{d,m} 2 (coef_lap(d,m) cos(q_d m h_d) + coef_nabk(d,m) sin(q_d m h_d)),
  which tends to |q + k|^2 / 2;
* the full hamiltonian with Bloch-phased projectors is Hermitian, and
  H(-k) = conj(H(k)) (time reversal);
* the Bloch-phased tables are the real tables times exp(-i k.x);
* sweeps in an empty lattice converge to the lowest closed-form eigenvalues.
"""
import math

import numpy as np
import pytest

from oracle.kpoint import KpointOracle
from rsdft_amd.system import Grid, Projectors, kinetic_coefficients, kpoint_kinetic, make_custom

K = np.array([0.11, -0.07, 0.05])


def _oracle(system, kvec, proj=None, **kw):
    ck2, nabk = kpoint_kinetic(system.Md, system.grid.h, kvec)
    pr = system.kpoint_projectors(kvec) if proj is None else proj
    return KpointOracle(system, pr, ck2, nabk, **kw)


@pytest.fixture(scope="module")
def small():
    # two atoms in a 16 x 12 x 12 cell: spheres cross the cell boundary
    pos = np.array([[0.3, 0.2, 0.1], [4.1, 3.0, 2.6]])
    return make_custom((16, 12, 12), (6.4, 5.2, 5.6), pos, Md=4, nband=6, seed=11)


def _empty(system):
    z = np.zeros(0, np.int32)
    return Projectors(0, z, np.zeros((0, 1), np.int32), np.zeros((0, 1), np.complex128), z, z, z,
                      z, z, np.zeros((0, 1), np.int32), np.zeros((0, 1, 3)), 0.0)


def test_k0_matches_real_oracle(small):
    from oracle.oracle import Oracle
    ro = Oracle(small)
    zo = _oracle(small, np.zeros(3))
    x = small.psi0()
    assert np.max(np.abs(zo.hamiltonian(x) - ro.hamiltonian(x))) < 1e-12 * np.max(np.abs(ro.hamiltonian(x)))
    x = ro.gram_schmidt(x)
    xr, er, rr = ro.conjugate_gradient(x)
    xz, ez, rz = zo.conjugate_gradient(x)
    assert np.max(np.abs(ez - er)) < 1e-11 and np.max(np.abs(xz.real - xr)) < 1e-9
    assert np.max(np.abs(xz.imag)) == 0.0
    g = zo.gram_schmidt(xz)
    assert np.max(np.abs(g.real - ro.gram_schmidt(xr))) < 1e-10
    _, wz, _ = zo.subspace_diag(g)
    _, wr, _ = ro.subspace_diag(ro.gram_schmidt(xr))
    assert np.max(np.abs(wz - wr)) < 1e-11
    for _ in range(2):
        xr, er, _ = ro.sweep(xr)
        xz, ez, _ = zo.sweep(xz)
    assert np.max(np.abs(ez - er)) < 1e-10


def test_kinetic_plane_waves_closed_form():
    N, L, Md = (12, 10, 14), (5.0, 4.4, 6.3), 4
    s = make_custom(N, L, np.zeros((0, 3)), Md=Md, nband=2)
    g = s.grid
    zo = _oracle(s, K, proj=_empty(s))
    c0, cl = kinetic_coefficients(Md, g.h)
    ck2, nabk = kpoint_kinetic(Md, g.h, K)
    x3, x2, x1 = np.meshgrid(*(np.arange(n) * h for n, h in zip(g.N[::-1], g.h[::-1])), indexing="ij")
    for n in ((1, 0, 0), (0, -2, 1), (3, 1, -2), (0, 0, 0)):
        q = [2 * math.pi * n[d] / g.L[d] for d in range(3)]
        psi = np.exp(1j * (q[0] * x1 + q[1] * x2 + q[2] * x3)).ravel()
        lam = c0 + ck2 + sum(2 * (cl[d, m] * math.cos(q[d] * (m + 1) * g.h[d])
                                  + nabk[d, m] * math.sin(q[d] * (m + 1) * g.h[d]))
                             for d in range(3) for m in range(Md))
        hp = zo.op_kinetic(psi)[0]
        assert np.max(np.abs(hp - lam * psi)) < 1e-11 * max(1.0, abs(lam))
        if max(abs(q[d]) * g.h[d] for d in range(3)) < 0.6:   # resolved by the stencil
            qk = sum((q[d] + K[d]) ** 2 for d in range(3)) / 2
            assert abs(lam - qk) < 1e-3 * max(qk, 0.1)


def test_hermitian_and_time_reversal(small):
    zo = _oracle(small, K)
    zm = _oracle(small, -K)
    rng = np.random.default_rng(2)
    ML = small.grid.ML
    a = rng.normal(size=(2, ML)) + 1j * rng.normal(size=(2, ML))
    ha = zo.hamiltonian(a)
    lhs = np.vdot(a[0], ha[1])
    rhs = np.conj(np.vdot(a[1], ha[0]))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    assert np.max(np.abs(zm.hamiltonian(np.conj(a)) - np.conj(ha))) < 1e-12 * np.max(np.abs(ha))


def test_bloch_tables_are_phased_real_tables(small):
    pr0 = small.proj
    prk = small.kpoint_projectors(K)
    h = np.array(small.grid.h)
    N = np.array(small.grid.N)
    assert np.array_equal(pr0.MJJ, prk.MJJ) and np.array_equal(pr0.JJP, prk.JJP)
    for l in range(pr0.nzlma):
        n = pr0.MJJ[l]
        a = pr0.amap[l] - 1
        j = pr0.JJP[l, :n] - 1
        idx = np.stack([j % N[0], (j // N[0]) % N[1], j // (N[0] * N[1])], axis=1)
        # unfold each point to the image nearest the atom
        R = small.positions[a]
        x = idx * h
        x = x - np.round((x - R) / (N * h)) * (N * h)
        want = pr0.uVk[l, :n] * np.exp(-1j * (x @ K))
        assert np.max(np.abs(prk.uVk[l, :n] - want)) < 1e-13


def test_empty_lattice_sweeps_converge():
    N, L, Md = (12, 12, 12), (4.0, 4.4, 4.8), 2
    s = make_custom(N, L, np.zeros((0, 3)), Md=Md, nband=5, seed=3)
    s.vloc[:] = 0.0
    zo = _oracle(s, K, proj=_empty(s))
    g = s.grid
    c0, cl = kinetic_coefficients(Md, g.h)
    ck2, nabk = kpoint_kinetic(Md, g.h, K)
    lam = []
    for n1 in range(-N[0] // 2, N[0] // 2):
        for n2 in range(-N[1] // 2, N[1] // 2):
            for n3 in range(-N[2] // 2, N[2] // 2):
                n = (n1, n2, n3)
                q = [2 * math.pi * n[d] / g.L[d] for d in range(3)]
                lam.append(c0 + ck2 + sum(2 * (cl[d, m] * math.cos(q[d] * (m + 1) * g.h[d])
                                               + nabk[d, m] * math.sin(q[d] * (m + 1) * g.h[d]))
                                          for d in range(3) for m in range(Md)))
    lam = np.sort(lam)[:5]
    x = zo.gram_schmidt(s.psi0c())
    for _ in range(25):
        x, e, _ = zo.sweep(x)
    assert np.max(np.abs(e - lam)) < 1e-8, (e, lam)
