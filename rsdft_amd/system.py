"""Synthetic Kohn-Sham inputs in RSDFT's own data layout (Gamma point, real psi).

Everything here is host-side set-up: the grid, the finite-difference
coefficients, a smooth random local potential, Kleinman-Bylander projector
tables laid out exactly as the reference lays them out, their position
derivatives, seeded initial wavefunctions and occupations.

Layout conventions (identical to the reference's Fortran arrays):

* A grid point (i1, i2, i3), 0 <= i_d < N_d, has the 1-D index
  ``i1 + i2*N1 + i3*N1*N2`` (x fastest), as in
  ``kinetic_sol_module.f90:85`` and ``ps_nloc2_module.f90:1171``.
* Wavefunctions are band-major: numpy shape ``(nb, ML)`` in C order is the
  memory of Fortran ``unk(ML, nb)``.
* Projector tables follow ``ps_nloc2_variables.f90``: ``JJP(MAXMJJ, nzlma)``
  (1-based grid index, ascending within a projector), ``uVk(MAXMJJ, nzlma)``,
  ``MJJ(nzlma)``, ``iuV(nzlma)``; numpy shape ``(nzlma, MAXMJJ)`` in C order is
  the Fortran column-major memory.  The projector order is atom-major, then
  orbital, then m = -L..L (``ps_nloc2_module.f90:449-455``).
* ``uV = sqrt(D) * f_l(r) * Y_lm`` with the real-Ylm convention of
  ``ylm_module.f90:10-72`` (for l = 1 and m = -1, 0, +1: -y, z, -x times
  sqrt(3/4pi)/r); ``iuV = sign(D) = +1`` (``ps_nloc2_module.f90:577``).
* Sphere test ``r^2 <= rc^2 + 1e-10`` (``ps_nloc2_module.f90:300``);
  duplicated points of a sphere that wraps onto itself are merged and exact
  zeros dropped (``ps_nloc2_module.f90:687-703, 1160-1182``); every list is
  sorted by grid index (``ps_nloc2_module.f90:923, 942-965``).

Random numbers come from a counter-based splitmix64 stream, so every array is
a pure function of (seed, stream, index) and can be regenerated piecewise.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Diamond 8-atom basis in fractional coordinates (input/input_Si8_atom:2-9).
DIAMOND_BASIS = np.array(
    [
        [0.00, 0.00, 0.00],
        [0.25, 0.25, 0.25],
        [0.50, 0.00, 0.50],
        [0.00, 0.50, 0.50],
        [0.50, 0.50, 0.00],
        [0.75, 0.25, 0.75],
        [0.25, 0.75, 0.75],
        [0.75, 0.75, 0.25],
    ]
)
CUBIC_A = 10.26          # bohr, cubic Si cell
POINTS_PER_CELL = 24     # grid points per cubic cell edge (h = 0.4275 bohr)
RC = 2.5                 # projector cut-off radius (bohr)
SIGMA = 1.2              # Gaussian width of r^l exp(-r^2/sigma^2)
VLOC_TERMS = 20
VLOC_AMPLITUDE = 0.5     # Ha per plane-wave term
VLOC_GMAX2 = 9           # |n|^2 <= 9, n integer, n != 0

STREAM_PSI = 1
STREAM_VLOC = 2
STREAM_D = 3
STREAM_PSI_IM = 4

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix_uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """Uniform [0,1) doubles, element k = splitmix64 finaliser of a counter.

    key_k = seed*0x9E3779B97F4A7C15 + stream*0xBF58476D1CE4E5B9 + offset + k
    (mod 2^64), then the splitmix64 mix; the top 53 bits scaled by 2^-53.
    """
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
                + np.uint64(stream) * np.uint64(0xBF58476D1CE4E5B9)
                + np.uint64(offset))
        z = base + np.arange(n, dtype=np.uint64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ----------------------------------------------------------------------------
# grid and finite-difference coefficients
# ----------------------------------------------------------------------------

@dataclasses.dataclass(frozen=True)
class Grid:
    N: tuple            # (N1, N2, N3)
    L: tuple            # cell edges (bohr), orthorhombic cell

    @property
    def h(self):
        """Hgrid = |a_d| / N_d (rgrid_sol_module.f90:71-73)."""
        return tuple(self.L[d] / self.N[d] for d in range(3))

    @property
    def dV(self):
        """dV = |det aa| / Ngrid(0) (rgrid_sol_module.f90:74)."""
        return (self.L[0] * self.L[1] * self.L[2]) / (self.N[0] * self.N[1] * self.N[2])

    @property
    def ML(self):
        return self.N[0] * self.N[1] * self.N[2]


def fd_lap_closed_form(Md: int) -> np.ndarray:
    """lap(0..Md) of the 2*Md-order central second derivative, closed form.

    lap(0) = -2 sum_{j=1..Md} 1/j^2,
    lap(j) = 2 (-1)^(j+1) (Md!)^2 / (j^2 (Md-j)! (Md+j)!).
    These are the weights ``get_coef_lapla_fd`` (fd_module.f90:18-54) builds
    from Lagrange products; the oracle restates those loops and is checked
    against this form.
    """
    lap = np.zeros(Md + 1)
    lap[0] = -2.0 * sum(1.0 / (j * j) for j in range(1, Md + 1))
    f = math.factorial
    for j in range(1, Md + 1):
        lap[j] = 2.0 * (-1) ** (j + 1) * f(Md) ** 2 / (j * j * f(Md - j) * f(Md + j))
    return lap


def kinetic_coefficients(Md: int, h) -> tuple[float, np.ndarray]:
    """coef_lap0 and coef_lap(3, Md) for an orthorhombic cell (ggg = 1,1,1).

    coef_lap(d, m) = lap(m) * (-0.5/h_d^2); coef_lap0 = lap(0) * sum_d(-0.5/h_d^2)
    (kinetic_module.f90:80-89).
    """
    lap = fd_lap_closed_form(Md)
    c = [-0.5 / (h[d] ** 2) for d in range(3)]
    coef_lap0 = lap[0] * (c[0] + c[1] + c[2])
    coef_lap = np.array([[lap[m] * c[d] for m in range(1, Md + 1)] for d in range(3)])
    return float(coef_lap0), coef_lap


def fd_nab_closed_form(Md: int) -> np.ndarray:
    """nab(1..Md) of the 2*Md-order central first derivative, closed form:
    nab(j) = (-1)^(j+1) (Md!)^2 / (j (Md-j)! (Md+j)!) (the weights
    get_coef_nabla_fd builds, fd_module.f90; nab(-j) = -nab(j))."""
    f = math.factorial
    return np.array([(-1) ** (j + 1) * f(Md) ** 2 / (j * f(Md - j) * f(Md + j))
                     for j in range(1, Md + 1)])


def kpoint_kinetic(Md: int, h, kvec) -> tuple[float, np.ndarray]:
    """const_k2 and coef_nabk(3, Md) of one k point, orthorhombic cell
    (kinetic_module.f90:118-149): coef_nabk(d, m) = nab(m)/h_d * k_d and
    const_k2 = |k|^2/2, k Cartesian (bohr^-1).  The stencil then applies
    zcoef(+m) = coef_lap - i coef_nabk to psi(x + m) and its conjugate to
    psi(x - m): -1/2 (grad + ik)^2 on the periodic part of a Bloch function."""
    nab = fd_nab_closed_form(Md)
    k = np.asarray(kvec, dtype=np.float64)
    coef_nabk = np.array([[nab[m] / h[d] * k[d] for m in range(Md)] for d in range(3)])
    return float(0.5 * np.dot(k, k)), coef_nabk


# ----------------------------------------------------------------------------
# local potential
# ----------------------------------------------------------------------------

def _vloc_wavevectors() -> np.ndarray:
    r = range(-3, 4)
    vecs = [(a, b, c) for a in r for b in r for c in r if 0 < a * a + b * b + c * c <= VLOC_GMAX2]
    return np.array(sorted(vecs), dtype=np.int64)


def make_vloc(grid: Grid, seed: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """V(r) = sum_j A cos(2 pi n_j . r / L + phi_j), 20 terms, A = 0.5 Ha.

    n_j are integer vectors with 0 < |n_j|^2 <= 9 chosen by the RNG from the
    sorted list of all such vectors; phi_j ~ U[0, 2 pi).  Returns
    (V[ML], n[20,3], phi[20]).
    """
    table = _vloc_wavevectors()
    u = splitmix_uniform(seed, STREAM_VLOC, 2 * VLOC_TERMS)
    pick = np.minimum((u[:VLOC_TERMS] * len(table)).astype(np.int64), len(table) - 1)
    nvec = table[pick]
    phi = 2.0 * math.pi * u[VLOC_TERMS:]
    N1, N2, N3 = grid.N
    i3, i2, i1 = np.meshgrid(np.arange(N3), np.arange(N2), np.arange(N1), indexing="ij")
    V = np.zeros((N3, N2, N1))
    for j in range(VLOC_TERMS):
        n = nvec[j]
        # 2 pi n.r/L with r_d = i_d h_d, evaluated from exact integer phases
        ph = (2.0 * math.pi) * ((n[0] * i1) / N1 + (n[1] * i2) / N2 + (n[2] * i3) / N3)
        V += VLOC_AMPLITUDE * np.cos(ph + phi[j])
    return V.reshape(-1), nvec, phi


# ----------------------------------------------------------------------------
# Kleinman-Bylander projectors
# ----------------------------------------------------------------------------

Y00 = 1.0 / math.sqrt(4.0 * math.pi)
K1 = math.sqrt(3.0 / (4.0 * math.pi))
# (l, m, Cartesian axis of r*Y_lm, sign) for the four channels, m = -L..L order
CHANNELS = ((0, 0, -1, 1.0), (1, -1, 1, -1.0), (1, 0, 2, 1.0), (1, 1, 0, -1.0))


def _sphere_offsets(h, rc):
    M = [int(math.ceil(rc / h[d])) + 2 for d in range(3)]
    d3, d2, d1 = np.meshgrid(np.arange(-M[2], M[2] + 1), np.arange(-M[1], M[1] + 1),
                             np.arange(-M[0], M[0] + 1), indexing="ij")
    return np.stack([d1.ravel(), d2.ravel(), d3.ravel()], axis=1)


def projector_values(x: np.ndarray, D: float, ch: int) -> tuple[np.ndarray, np.ndarray]:
    """uV and duV/dR at displacement vectors x (P,3) = r - R, one channel.

    uV = sqrt(D) f_l(|x|) Y_lm(x) with f_l = r^l exp(-r^2/sigma^2); for l = 1
    f_1 Y_1m = sqrt(D) * sign * sqrt(3/4pi) * x_axis * exp(-r^2/sigma^2).
    duV/dR = -grad_r uV (the atom moves, the grid point does not).
    """
    l, m, axis, sgn = CHANNELS[ch]
    r2 = np.sum(x * x, axis=1)
    g = np.exp(-r2 / (SIGMA * SIGMA))
    sD = math.sqrt(D)
    if l == 0:
        c = sD * Y00
        val = c * g
        grad = (c * g * (-2.0 / SIGMA ** 2))[:, None] * x
    else:
        c = sD * sgn * K1
        xa = x[:, axis]
        val = c * xa * g
        grad = (-2.0 / SIGMA ** 2) * (c * xa * g)[:, None] * x
        grad[:, axis] += c * g
    return val, -grad


@dataclasses.dataclass
class Projectors:
    """KB tables in the reference layout plus derivative tables for forces."""
    nzlma: int
    MJJ: np.ndarray          # (nzlma,) int32
    JJP: np.ndarray          # (nzlma, MAXMJJ) int32, 1-based
    uVk: np.ndarray          # (nzlma, MAXMJJ) float64
    iuV: np.ndarray          # (nzlma,) int32
    amap: np.ndarray         # (nzlma,) int32, 1-based atom index
    lmap: np.ndarray
    mmap: np.ndarray
    MJJ_MAP: np.ndarray      # (nzlma,) full sphere point counts (zeros kept)
    JJ_MAP: np.ndarray       # (nzlma, MMJJ) 1-based grid index of every sphere point
    duVdR: np.ndarray        # (nzlma, MMJJ, 3)  Fortran duVdR(3, MMJJ, nzlma)
    min_boundary_gap: float  # min over points of |rc - r| (bohr)

    @property
    def MAXMJJ(self):
        return self.JJP.shape[1]

    @property
    def MMJJ(self):
        return self.JJ_MAP.shape[1]


def make_projectors(grid: Grid, atom_ic: np.ndarray, atom_delta: np.ndarray,
                    D: np.ndarray, sphere_delta: np.ndarray | None = None,
                    kvec=None) -> Projectors:
    """Build uV/JJP/MJJ tables for atoms at ic*h + delta (ic integer grid index).

    ``sphere_delta`` (default: ``atom_delta``) is the atom offset used for the
    sphere membership test only; passing the unmoved offsets freezes every
    point set while the values follow the moved atom, which is the function
    whose R-derivative the nonlocal force is.

    With ``kvec`` (Cartesian, bohr^-1) the values are Bloch-phased,
    uVk = uV exp(-i k.x) with x the unfolded position of the point (the image
    it came from), summed over the images of one grid point
    (prep_uvk_ps_nloc2, ps_nloc2_module.f90:1128-1190): complex tables.
    """
    h = np.array(grid.h)
    N = np.array(grid.N)
    offs = _sphere_offsets(grid.h, RC)
    natom = atom_ic.shape[0]
    nz = 4 * natom
    lists, full = [], []
    gap = np.inf
    if sphere_delta is None:
        sphere_delta = atom_delta
    for a in range(natom):
        x = offs * h[None, :] - atom_delta[a][None, :]
        xt = offs * h[None, :] - sphere_delta[a][None, :]
        r2 = np.sum(xt * xt, axis=1)
        keep = r2 <= RC * RC + 1.0e-10
        xs = x[keep]
        gp = np.mod(atom_ic[a][None, :] + offs[keep], N[None, :])
        idx = gp[:, 0] + N[0] * (gp[:, 1] + N[1] * gp[:, 2])
        r_in = np.sqrt(r2[keep])
        r_out = np.sqrt(r2[~keep])
        gap = min(gap, float(np.min(RC - r_in)), float(np.min(r_out - RC)))
        if kvec is not None:
            xu = (atom_ic[a][None, :] + offs[keep]) * h[None, :]
            phase = np.exp(-1j * (xu @ np.asarray(kvec, dtype=np.float64)))
        for ch in range(4):
            val, dval = projector_values(xs, float(D[a, ch]), ch)
            if kvec is not None:
                val = val * phase
            order = np.argsort(idx, kind="stable")
            si, sv, sd = idx[order], val[order], dval[order]
            full.append((si + 1, sd))
            uniq, inv = np.unique(si, return_inverse=True)
            summed = np.zeros(len(uniq), dtype=sv.dtype)
            np.add.at(summed, inv, sv)
            nonzero = np.zeros(len(uniq), dtype=bool)
            np.logical_or.at(nonzero, inv, sv != 0.0)
            lists.append((uniq[nonzero] + 1, summed[nonzero]))
    if not lists:                       # no atoms: an empty operator
        z = np.zeros(0, np.int32)
        return Projectors(0, z, np.zeros((0, 1), np.int32),
                          np.zeros((0, 1), np.complex128 if kvec is not None else np.float64),
                          z, z, z, z, z, np.zeros((0, 1), np.int32), np.zeros((0, 1, 3)), 0.0)
    maxmjj = max(len(li[0]) for li in lists)
    mmjj = max(len(f[0]) for f in full)
    MJJ = np.zeros(nz, np.int32)
    JJP = np.zeros((nz, maxmjj), np.int32)
    uVk = np.zeros((nz, maxmjj), dtype=np.complex128 if kvec is not None else np.float64)
    MJJ_MAP = np.zeros(nz, np.int32)
    JJ_MAP = np.zeros((nz, mmjj), np.int32)
    duVdR = np.zeros((nz, mmjj, 3))
    for k, (ji, jv) in enumerate(lists):
        MJJ[k] = len(ji)
        JJP[k, :len(ji)] = ji
        uVk[k, :len(ji)] = jv
    for k, (fi, fd) in enumerate(full):
        MJJ_MAP[k] = len(fi)
        JJ_MAP[k, :len(fi)] = fi
        duVdR[k, :len(fi)] = fd
    amap = np.repeat(np.arange(1, natom + 1, dtype=np.int32), 4)
    lmap = np.tile(np.array([c[0] for c in CHANNELS], np.int32), natom)
    mmap = np.tile(np.array([c[1] for c in CHANNELS], np.int32), natom)
    return Projectors(nz, MJJ, JJP, uVk, np.ones(nz, np.int32), amap, lmap, mmap,
                      MJJ_MAP, JJ_MAP, duVdR, gap)


# ----------------------------------------------------------------------------
# whole system
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class System:
    name: str
    grid: Grid
    Md: int
    atom_ic: np.ndarray      # (natom, 3) int nearest-grid index
    atom_delta: np.ndarray   # (natom, 3) bohr offset from that grid point
    D: np.ndarray            # (natom, 4)
    vloc: np.ndarray         # (ML,)
    proj: Projectors
    nband: int
    occ: np.ndarray          # (nband,)
    seed: int
    coef_lap0: float
    coef_lap: np.ndarray     # (3, Md)

    @property
    def natom(self):
        return self.atom_ic.shape[0]

    @property
    def positions(self):
        return self.atom_ic * np.array(self.grid.h)[None, :] + self.atom_delta

    def psi0(self, band0: int = 0, nbands: int | None = None, dtype=np.float64) -> np.ndarray:
        """Initial psi, uniform(-1, 1) per element; shape (nbands, ML)."""
        nb = self.nband - band0 if nbands is None else nbands
        ML = self.grid.ML
        u = splitmix_uniform(self.seed, STREAM_PSI, nb * ML, offset=band0 * ML)
        return (2.0 * u - 1.0).reshape(nb, ML).astype(dtype, copy=False)

    def psi0c(self, nbands: int | None = None) -> np.ndarray:
        """Complex start for a k point: real part psi0(), imaginary part
        uniform(-1, 1) from its own stream; shape (nbands, ML)."""
        nb = self.nband if nbands is None else nbands
        ML = self.grid.ML
        v = splitmix_uniform(self.seed, STREAM_PSI_IM, nb * ML)
        return self.psi0(0, nb) + 1j * (2.0 * v - 1.0).reshape(nb, ML)

    def kpoint_projectors(self, kvec) -> Projectors:
        """Bloch-phased tables of this system's projectors at k (complex uVk)."""
        return make_projectors(self.grid, self.atom_ic, self.atom_delta, self.D, kvec=kvec)


CONFIGS = {
    # name: (cubic cells per edge, bands, displaced atom delta or None)
    "si8": (1, 32, None),
    "si216": (3, 480, None),
    "si64": (2, 160, (0.1, 0.05, 0.0)),
    "si1000": (5, 2200, None),
}


def make_system(ncell: int, nband: int, seed: int = 20240601, Md: int = 6,
                displace=None, name: str | None = None, points_per_cell: int = POINTS_PER_CELL,
                nocc: int | None = None) -> System:
    """Diamond Si supercell of ncell^3 cubic cells on an N = 24*ncell grid.

    ``displace`` moves atom 0 by the given Cartesian vector (bohr).
    """
    L = CUBIC_A * ncell
    Nside = points_per_cell * ncell
    grid = Grid((Nside,) * 3, (L,) * 3)
    frac = []
    for k3 in range(ncell):
        for k2 in range(ncell):
            for k1 in range(ncell):
                for b in DIAMOND_BASIS:
                    frac.append((b + np.array([k1, k2, k3])) / ncell)
    frac = np.array(frac)
    icf = frac * Nside
    ic = np.rint(icf).astype(np.int64)
    if np.max(np.abs(icf - ic)) > 1e-9:
        raise ValueError("atoms must sit on grid points before displacement")
    delta = np.zeros((len(ic), 3))
    if displace is not None:
        delta[0] = np.asarray(displace, dtype=np.float64)
    natom = len(ic)
    D = 0.5 + 1.5 * splitmix_uniform(seed, STREAM_D, natom * 4).reshape(natom, 4)
    vloc, _, _ = make_vloc(grid, seed)
    proj = make_projectors(grid, ic, delta, D)
    if nocc is None:
        nocc = min(2 * natom, nband)     # 4 valence electrons per Si, 2 per band
    occ = np.zeros(nband)
    occ[:nocc] = 2.0
    c0, cl = kinetic_coefficients(Md, grid.h)
    return System(name or f"si{natom}", grid, Md, ic, delta, D, vloc, proj, nband, occ,
                  seed, c0, cl)


def make_config(name: str, seed: int = 20240601, nband: int | None = None) -> System:
    ncell, nb, disp = CONFIGS[name]
    return make_system(ncell, nb if nband is None else nband, seed=seed, displace=disp, name=name)


def make_custom(N, L, positions, Md: int = 6, nband: int = 8, seed: int = 7,
                nocc: int | None = None, name: str = "custom", D=None) -> System:
    """System on an orthorhombic N[0]xN[1]xN[2] grid of a cell with edges L,
    atoms at Cartesian ``positions`` (bohr, shape (natom, 3))."""
    grid = Grid(tuple(int(n) for n in N), tuple(float(x) for x in L))
    h = np.array(grid.h)
    pos = np.atleast_2d(np.asarray(positions, dtype=np.float64))
    ic = np.rint(pos / h[None, :]).astype(np.int64)
    delta = pos - ic * h[None, :]
    ic = np.mod(ic, np.array(grid.N)[None, :])
    natom = len(ic)
    if D is None:
        D = 0.5 + 1.5 * splitmix_uniform(seed, STREAM_D, natom * 4).reshape(natom, 4)
    vloc, _, _ = make_vloc(grid, seed)
    proj = make_projectors(grid, ic, delta, D)
    if nocc is None:
        nocc = min(2 * natom, nband)
    occ = np.zeros(nband)
    occ[:nocc] = 2.0
    c0, cl = kinetic_coefficients(Md, grid.h)
    return System(name, grid, Md, ic, delta, np.asarray(D), vloc, proj, nband, occ, seed, c0, cl)
