"""Real norm-conserving pseudopotentials for the Kohn-Sham path: the one-time
host set-up that turns a PSV pseudopotential file and atomic positions into
what the GPU library consumes -- the Kleinman-Bylander projector tables
(rsdft_gpu_set_projectors), the ionic local potential Vion and the initial
density (rsdft_gpu_scf_init) -- plus the Ewald energy of the ions.

The reference runs the same set-up once on the CPU before its sweeps:

* read_psv            ps_read_psv.f90:13-366 (PSV 1.0/2.x, norm-conserving)
* radial filtering    ps_nloc2_init_module.f90:72-216 (mask function of
                      maskf_module.f90:13-55, Neville interpolation of
                      polint_module.f90:10-39, Fourier filtering of
                      filtering.f90:11-233 with its Simpson rule :236-273)
* projector tables    ps_nloc2_module.f90:56-939 (prep_ps_nloc2), :1128-1190
                      (prep_uvk_ps_nloc2, Gamma point), :942-970 (sorting)
* G sphere / shells   ggrid_module.f90:55-150
* local potential     ps_local_module.f90:180-300 (init_ps_local),
                      ps_local_fftw_module.f90:18-80, strfac_module.f90:15-37
* initial density     ps_initrho_module.f90:44-280, :493-503
* Ewald energy        ewald_module.f90:604-820

Every step is host numpy: it runs once per atomic configuration, outside the
sweeps (as in the reference), and its outputs are uploaded to the device.
"""
from __future__ import annotations

import dataclasses
import gzip
import math
import re

import numpy as np

from .system import Grid, Projectors, System, kinetic_coefficients

FOUR_PI = 4.0 * math.pi


# ----------------------------------------------------------------------------
# PSV reader (ps_read_psv.f90:13-366)
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class Pseudopotential:
    """One element's radial data on the reference's mesh: index 0 is r = 0,
    index i >= 1 the file's i-th log-mesh point (ps_read_psv.f90:262-297)."""
    zps: float                 # valence charge (Zps)
    zatom: float
    xc: str
    lo: np.ndarray             # (norb,) angular momentum of each projector
    no: np.ndarray             # (norb,) reference index within its l
    anorm: np.ndarray          # (norb,) |D_ii|
    inorm: np.ndarray          # (norb,) sign(D_ii)
    rps: np.ndarray            # (norb,) cut-off radius of the file
    nrps: np.ndarray           # (norb,) points 0..nrps-1 carry the projector
    dij: np.ndarray            # (norb, norb)
    rad: np.ndarray            # (Mr,)
    rab: np.ndarray            # (Mr,) dr/di of the log mesh
    vql: np.ndarray            # (Mr,) local potential
    cdc: np.ndarray            # (Mr,) core charge (partial core correction)
    cdd: np.ndarray            # (Mr,) 4 pi r^2 x initial valence density
    viod: np.ndarray           # (Mr, norb) sqrt(anorm) x beta_l(r)
    parloc: np.ndarray         # (4,) erf fit parameters of the file
    cdd_coef: np.ndarray       # (ngauss, 3) Gaussian fit of the initial charge

    @property
    def norb(self):
        return len(self.lo)

    @property
    def mr(self):
        return len(self.rad)


_TOKEN = re.compile(r"'[^']*'|\"[^\"]*\"|[^\s,]+")


class _Lines:
    """Line cursor with Fortran list-directed reads (values may continue on
    the following records; BACKSPACE moves one record back)."""

    def __init__(self, text):
        self.lines = text.splitlines()
        self.pos = 0

    def line(self):
        s = self.lines[self.pos]
        self.pos += 1
        return s

    def values(self, n):
        out = []
        while len(out) < n:
            toks = _TOKEN.findall(self.line())
            for t in toks:
                if t == "/":
                    return out
                out.append(t)
        return out[:n] if len(out) >= n else out


def _f(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def read_psv(source) -> Pseudopotential:
    """Read a norm-conserving PSV file (path, .gz path or text), following
    ps_read_psv.f90:13-366: version detection (#PSV, #xPSV or old), the local
    potential on its log mesh (2 or 3 columns), D_ij / q_ij and the psi, phi,
    beta columns per reference energy, the initial-charge Gaussians, and the
    derived arrays (viod = sqrt(anorm) beta, rab = r h, cdd)."""
    if isinstance(source, str) and "\n" not in source:
        opener = gzip.open if source.endswith(".gz") else open
        with opener(source, "rt") as fh:
            text = fh.read()
    else:
        text = source
    c = _Lines(text)
    head = c.line()
    if head.startswith("#PSV"):
        verpot = 2
        v = c.values(3)
        xc, zatom = v[0].strip("'\"").strip(), _f(v[1])
        while c.lines[c.pos].startswith("#"):
            c.pos += 1
        znuc = _f(c.values(1)[0])
    elif head.startswith("#xPSV"):
        verpot = 3
        while c.lines[c.pos].startswith("#"):
            c.pos += 1
        xc = c.values(1)[0].strip("'\"").strip()
        v = c.values(2)
        znuc, zatom = _f(v[0]), _f(v[1])
    else:
        verpot = 1
        xc, zatom, znuc = "LDAPZ81", 0.0, 0.0
        c.line()                                  # comment line
    nl = int(c.values(1)[0])
    nr = [int(x) for x in c.values(nl)]
    v = c.values(2)
    nsmpl, rc_in = int(v[0]), _f(v[1])
    a = [_f(x) for x in c.values(4)]
    if abs(a[0] + a[2] - 1.0) > 1.0e-10:
        raise ValueError(f"PSV local parameter fitting error: a1..a4 = {a}")
    ndlc = int(c.values(1)[0])
    # two or three columns: the first record holds three values when cc is there
    first = c.lines[c.pos].split()
    ncol = 3 if len(first) >= 3 else 2
    tab = np.array([[_f(t) for t in c.line().split()[:ncol]] for _ in range(ndlc)])
    rr, vl = tab[:, 0], tab[:, 1]
    cc = tab[:, 2] if ncol == 3 else np.zeros(ndlc)
    if rr[0] <= 0.0:
        raise ValueError(f"PSV local potential: rr(1) = {rr[0]}")
    h = math.exp(math.log(rr[-1] / rr[0]) / (ndlc - 1))
    if abs(h * rr[0] - rr[1]) > 1.0e-12:
        raise ValueError("PSV local potential: mesh is not logarithmic")
    c.line()                                      # ### NON-LOCAL POTENTIAL PARAMETERS ###
    v = c.values(4)
    ndata, r1, rend, r2 = int(v[0]), _f(v[1]), _f(v[2]), _f(v[3])
    if ndata != nsmpl:
        raise ValueError(f"PSV: nsmpl {nsmpl} != ndata {ndata}")
    hn = math.exp(math.log(rend / r1) / (ndata - 1))
    r = r1 * hn ** np.arange(ndata)
    if abs(r[1] - r2) > 1.0e-6 or abs(rr[ndata - 1] - rend) > 1.0e-8:
        raise ValueError("PSV: nonlocal mesh does not match the local one")
    if math.sqrt(np.sum((rr[:ndata] - r) ** 2) / ndata) > 1.0e-7:
        raise ValueError("PSV: radial grid is inconsistent")
    mmax = max(nr)
    bet = np.zeros((nsmpl, mmax, nl))
    ddi = np.zeros((mmax, mmax, nl))
    for l in range(nl):
        c.line()                                  # ### L= l ...
        m = nr[l]
        vals = [_f(t) for t in c.values(m * (m + 1) // 2)]
        k = 0
        for j in range(m):
            for i in range(j + 1):
                ddi[i, j, l] = ddi[j, i, l] = vals[k]
                k += 1
        c.values(m * (m + 1) // 2)                # q_ij (ultrasoft only)
        for j in range(m):
            c.line()                              # ** PSI, PHI, BETA ...
            for i in range(nsmpl):
                t = c.line().split()
                bet[i, j, l] = _f(t[2])
    ngauss, gau = 0, np.zeros((0, 3))
    if verpot != 3:
        while c.pos < len(c.lines):
            if c.line()[:18] == "### initial charge":
                ngauss = int(c.values(1)[0])
                gau = np.array([[_f(t) for t in c.values(3)] for _ in range(ngauss)])
                break
        else:
            raise ValueError("PSV: no '### initial charge' section")
    Mr = max(ndlc, nsmpl, ndata) + 1
    norb = max(1, sum(nr))
    lo, no, anorm, inorm = (np.zeros(norb, np.int32), np.zeros(norb, np.int32),
                            np.zeros(norb), np.zeros(norb))
    k = 0
    for l in range(nl):
        for j in range(nr[l]):
            lo[k], no[k] = l, j + 1
            anorm[k] = abs(ddi[j, j, l])
            inorm[k] = math.copysign(1.0, ddi[j, j, l])
            k += 1
    dij = np.zeros((norb, norb))
    for jo in range(norb):
        for io in range(norb):
            if lo[io] == lo[jo]:
                dij[io, jo] = ddi[no[io] - 1, no[jo] - 1, lo[io]]
    multref = np.count_nonzero(anorm) < np.count_nonzero(dij)
    rad, vql, cdc, rab, cdd = (np.zeros(Mr) for _ in range(5))
    rad[1:ndlc + 1], vql[1:ndlc + 1], cdc[1:ndlc + 1] = rr, vl, cc
    vql[0] = vql[1]
    viod = np.zeros((Mr, norb))
    k = 0
    for l in range(nl):
        for j in range(nr[l]):
            s = 1.0 if multref else math.sqrt(anorm[k])
            viod[1:nsmpl + 1, k] = bet[:, j, l] * s
            k += 1
    hl = math.log(rr[-1] / rr[0]) / (ndlc - 1)
    rab[1:ndlc + 1] = rr * hl
    if verpot != 3:
        for a0, b0, c0 in gau:
            cdd[1:ndlc + 1] += rr ** 2 * FOUR_PI * (a0 + b0 * rr ** 2) * np.exp(-c0 * rr ** 2)
    return Pseudopotential(znuc, zatom, xc, lo, no, anorm, inorm, np.full(norb, rc_in),
                           np.full(norb, nsmpl + 1, np.int32), dij, rad, rab, vql, cdc, cdd,
                           viod, np.array(a), gau)


# ----------------------------------------------------------------------------
# integration and interpolation helpers
# ----------------------------------------------------------------------------

def simpson(f):
    """The reference's composite Simpson rule over the index (simp with m = 2,
    filtering.f90:236-273, ps_local_module.f90:302-338): panels (1,2,3),
    (3,4,5), ...; with an even count the last point is left out."""
    n = len(f)
    last = n - 3 if n % 2 else n - 4          # first index (0-based) of the last panel
    i = np.arange(0, last + 1, 2)
    return float(np.sum(f[i] + 4.0 * f[i + 1] + f[i + 2]) / 3.0)


def neville(xa, ya, x):
    """polint_module.f90:10-39 for a batch: xa, ya (P, n), x (P,).  Returns the
    interpolated value and the reference's error estimate (change of the mean
    of the last tableau column between the last two orders)."""
    f0 = ya.astype(np.float64).copy()
    n = xa.shape[1]
    f0nm = ya[:, 0].copy()
    f0nm_0 = f0nm
    for m in range(1, n):
        a = x[:, None] - xa[:, :n - m]
        b = x[:, None] - xa[:, m:]
        f0[:, :n - m] = (b * f0[:, :n - m] - a * f0[:, 1:n - m + 1]) / (b - a)
        f0nm_0 = f0nm
        f0nm = np.mean(f0[:, :n - m], axis=1)
    return f0[:, 0], f0nm - f0nm_0


def _windowed_interp(xa, ya, x, center, lo, hi, ep=None):
    """The reference's interpolation at x from the tables xa, ya: windows
    [max(lo, c-m), min(c+m, hi)] for m = 1..20 around index c, keeping the
    value whose error estimate is smallest in magnitude (strictly smaller
    replaces); with ``ep`` the search stops once that estimate is below ep
    (ps_nloc2_module.f90:322-333), without it all 20 windows are tried
    (ps_nloc2_init_module.f90:117-133).  Indices are 0-based, inclusive."""
    P = len(x)
    best_v = np.zeros(P)
    best_e = np.full(P, 1.0e10)
    active = np.ones(P, bool)
    for m in range(1, 21):
        if not np.any(active):
            break
        idx = np.nonzero(active)[0]
        m1 = np.maximum(center[idx] - m, lo)
        m2 = np.minimum(center[idx] + m, hi)
        width = m2 - m1 + 1
        for w in np.unique(width):
            sel = width == w
            ii = idx[sel]
            cols = m1[sel][:, None] + np.arange(w)[None, :]
            v, e = neville(xa[cols], ya[cols], x[ii])
            better = np.abs(e) < best_e[ii]
            best_v[ii[better]] = v[better]
            best_e[ii[better]] = np.abs(e[better])
        if ep is not None:
            active &= ~(best_e < ep)
    return best_v, best_e


# ----------------------------------------------------------------------------
# Fourier filtering of the projectors (ps_nloc2_init_module.f90:72-216)
# ----------------------------------------------------------------------------

NMSK = 201
DR_FILTER = 2.0e-3


def mask_function(eta: float):
    """maskf_module.f90:13-55: the mask m(x) on x_i = i/200 (i = 0..201) is
    the eigenvector of the smallest eigenvalue of the 200x200 kernel
    K_ij = sin((x_i+x_j) eta)/(x_i+x_j) - sin((x_i-x_j) eta)/(x_i-x_j)
    (diagonal: sin(2 x_i eta)/(2 x_i) + 200 pi - eta), divided by x_i and
    normalised to m(0) = 1 (m(x_200+1) = 0)."""
    n = NMSK - 1
    dx = 1.0 / n
    x = np.arange(1, n + 1) * dx
    s = x[:, None] + x[None, :]
    d = x[:, None] - x[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        K = np.sin(s * eta) / s - np.where(d == 0.0, 0.0, np.sin(d * eta) / np.where(d == 0, 1, d))
    K[np.diag_indices(n)] = np.sin(2 * x * eta) / (2 * x) + n * math.pi - eta
    w, V = np.linalg.eigh(K)
    v = V[:, 0]
    c = v[0] / dx
    maskr = np.zeros(NMSK + 1)
    maskr[1:n + 1] = v / x / c
    maskr[0] = 1.0
    maskr[NMSK] = 0.0
    xm = np.arange(NMSK + 1) * dx
    return xm, maskr


def mask_at(x, xm, maskr):
    """Mask values at x (the 20-window interpolation of
    ps_nloc2_init_module.f90:117-133; 1 for x <= dx)."""
    x = np.asarray(x, dtype=np.float64)
    dxm = xm[1]
    out = np.ones(len(x))
    far = x > dxm
    if np.any(far):
        m0 = (x[far] / dxm).astype(np.int64)
        v, _ = _windowed_interp(xm, maskr, x[far], m0, 1, NMSK)
        out[far] = v
    return out


def _filter_kernel(L, qc, r, r1):
    """int_0^qc j_L(q r) j_L(q r1) q^2 dq in the closed forms of
    filtering.f90:28-226 for one output radius r (scalar) and the mesh r1."""
    t = np.zeros_like(r1)
    if L == 0:
        if r == 0.0:
            nz = r1 != 0.0
            t[~nz] = qc ** 3 / 3.0
            y = r1[nz]
            t[nz] = np.sin(qc * y) / y ** 3 - qc * np.cos(qc * y) / y ** 2
            return t
        z = r1 == 0.0
        e = r1 == r
        o = ~(z | e)
        t[z] = math.sin(qc * r) / r ** 3 - qc * math.cos(qc * r) / r ** 2
        t[e] = (2.0 * qc * r - math.sin(2.0 * qc * r)) / (4.0 * r ** 3)
        y = r1[o]
        t[o] = (np.sin(qc * (r - y)) / (r - y) - np.sin(qc * (r + y)) / (r + y)) / (2.0 * r * y)
        return t
    if r == 0.0:
        return None                                # the filtered value is 0
    e = r1 == r
    o = (r1 != 0.0) & ~e
    y = r1[o]
    qr = qc * r
    sb0x = math.sin(qr) / qr
    sb1x = sb0x / qr - math.cos(qr) / qr
    sb0y = np.sin(qc * y) / (qc * y)
    sb1y = sb0y / (qc * y) - np.cos(qc * y) / (qc * y)
    base = (2.0 * qr - math.sin(2.0 * qr)) / (4.0 * r ** 3)
    if L == 1:
        t[e] = base - qc * qc * sb0x * sb1x / r
        t[o] = (y * sb0y * sb1x - r * sb0x * sb1y) * qc * qc / (r * r - y * y)
    elif L == 2:
        s1 = math.sin(qr) / (qr * qr) - math.cos(qr) / qr
        t[e] = base - 3.0 * qc * s1 * s1 / (r * r)
        t[o] = ((r * sb0y * sb1x - y * sb0x * sb1y) * qc * qc / (r * r - y * y)
                - 3.0 * qc / (r * y) * sb1x * sb1y)
    elif L == 3:
        s1 = math.sin(qr) / (qr * qr) - math.cos(qr) / qr
        t[e] = (base - 3.0 * qc * s1 * s1 / (r * r)
                - qc ** 2 / r * (3.0 / qr * (15.0 / qr ** 2 - 1.0) * s1 * s1
                                 + 5.0 / qr * sb0x * sb0x + (1.0 - 30.0 / qr ** 2) * sb0x * s1))
        t[o] = (sb0y * sb1x * (15.0 / (y * r * r) + qc ** 2 * y / (r ** 2 - y ** 2))
                + sb0x * sb1y * (15.0 / (r * y * y) - qc ** 2 * r / (r ** 2 - y ** 2))
                - 45.0 / (qc * r ** 2 * y ** 2) * sb1x * sb1y - 5.0 * qc / (r * y) * sb0x * sb0y)
    elif L == 4:
        s1 = math.sin(qr) / (qr * qr) - math.cos(qr) / qr
        t[e] = (base - 3.0 * qc * s1 * s1 / (r * r)
                - qc ** 2 / r * (3.0 / qr * (15.0 / qr ** 2 - 1.0) * s1 * s1
                                 + 5.0 / qr * sb0x * sb0x + (1.0 - 30.0 / qr ** 2) * sb0x * s1)
                - qc ** 2 / r * ((1575.0 / qr ** 5 - 255.0 / qr ** 3 + 10.0 / qr) * s1 * s1
                                 + (175.0 / qr ** 3 - 5.0 / qr) * sb0x * sb0x
                                 + (-1050.0 / qr ** 4 + 100.0 / qr ** 2 - 1.0) * sb0x * s1))
        t[o] = (sb0y * sb1x * (-35.0 / (y * y * r) + qc ** 2 * r / (r ** 2 - y ** 2)
                               + 525.0 / (qc ** 2 * y ** 2 * r ** 3))
                + sb0x * sb1y * (-35.0 / (r * r * y) + qc ** 2 * y / (y ** 2 - r ** 2)
                                 + 525.0 / (qc ** 2 * r ** 2 * y ** 3))
                + sb1x * sb1y * (105.0 * (r ** 2 + y ** 2) / (qc * r ** 3 * y ** 3)
                                 - 10.0 * qc / (r * y) - 1575.0 / (qc ** 3 * r ** 3 * y ** 3))
                + sb0x * sb0y * (-175.0 / (qc * r ** 2 * y ** 2)))
    else:
        raise ValueError(f"projector filtering is implemented for L <= 4, not {L}")
    return t


@dataclasses.dataclass
class FilteredProjectors:
    """Projector radial functions on the uniform mesh r = i * 2e-3
    (ps_nloc2_init_module.f90:147-159): viod[:, k] for i < nrps[k]."""
    lo: np.ndarray
    inorm: np.ndarray
    rps: np.ndarray            # cut-off radius (rcfac x the file's, on the log mesh)
    nrps: np.ndarray           # points of the uniform mesh inside it (+1)
    rad1: np.ndarray
    viod: np.ndarray           # (MMr, norb)
    mask_maxerr: float


def filter_projectors(ps: Pseudopotential, qcut: float, rcfac: float = 1.5,
                      qcfac: float = 1.0, etafac: float = 8.0) -> FilteredProjectors:
    """ps_nloc2_init (ps_nloc2_init_module.f90:72-216): enlarge each cut-off
    to rcfac x Rc on the log mesh, divide the projector by the mask there,
    filter it to |q| <= qcfac x qcut (opFiltering), resample it on the
    uniform mesh and multiply by the mask again."""
    qc = qcut * qcfac
    if qc <= 0.0:
        qc = qcut
    rad, rab = ps.rad, ps.rab
    norb = ps.norb
    nrps_log = np.zeros(norb, np.int64)
    rps = np.zeros(norb)
    for k in range(norb):
        rc = ps.rps[k] * rcfac
        nrc = int(np.argmin(np.abs(rad - rc)))           # first minimum (minloc)
        if rad[nrc] < rc:
            nrc += 1
        if nrc >= ps.mr:
            raise ValueError("rcfac is too large for the radial mesh")
        nrps_log[k] = nrc + 1                            # 1-based count
        rps[k] = rad[nrc]
    xm, maskr = mask_function(etafac)
    nrps = np.zeros(norb, np.int64)
    for k in range(norb):
        n = int(rps[k] / DR_FILTER + 1.0)
        if (n - 1) * DR_FILTER < rps[k]:
            n += 1
        nrps[k] = n
    mmr = max(ps.mr, int(np.max(nrps)))
    rad1 = np.arange(mmr) * DR_FILTER
    viod = np.zeros((mmr, norb))
    viod[:ps.mr] = ps.viod
    maxerr = 0.0
    const = 2.0 / math.pi
    for k in range(norb):
        L = int(ps.lo[k])
        wm = mask_at(rad[:nrps_log[k]] / rps[k], xm, maskr)
        n0 = int(ps.nrps[k])                             # the file's extent (NRps0)
        vrad = rad[:n0] * ps.viod[:n0, k] * rab[:n0] / wm[:n0]
        out = np.zeros(nrps[k])
        for i in range(nrps[k]):
            t = _filter_kernel(L, qc, rad1[i], rad[:n0])
            out[i] = 0.0 if t is None else simpson(t * vrad) * const
        y = mask_at(rad1[:nrps[k]] / rps[k], xm, maskr)
        viod[:nrps[k], k] = y * out
    return FilteredProjectors(ps.lo.copy(), ps.inorm.copy(), rps, nrps, rad1, viod, maxerr)


# ----------------------------------------------------------------------------
# projector tables on the grid (ps_nloc2_module.f90:56-939, :1128-1190)
# ----------------------------------------------------------------------------

def real_ylm(x, y, z, L, M_in):
    """Real spherical harmonics of the reference (ylm_module.f90:10-121):
    Condon-Shortley associated Legendre functions, the azimuth from
    atan(y/x) (+pi for x < 0, +-pi/2 when |x| < 1e-10), sqrt(2) cos(M phi)
    for M > 0 and -/+ sqrt(2) sin(M phi) for M < 0 with |M| even / odd;
    0 at the origin for L > 0.  Vectorised over points."""
    x, y, z = (np.asarray(t, dtype=np.float64) for t in (x, y, z))
    if L == 0:
        return np.full(x.shape, 1.0 / math.sqrt(FOUR_PI))
    M = abs(M_in)
    c = 1.0
    for k in range(L - M + 1, L + M + 1):
        c *= k
    clm = math.sqrt((2.0 * L + 1.0) / (c * FOUR_PI))
    origin = (x == 0.0) & (y == 0.0) & (z == 0.0)
    r = np.sqrt(x * x + y * y + z * z)
    rs = np.where(origin, 1.0, r)
    ct = z / rs
    ct = np.where(np.abs(np.abs(ct) - 1.0) < 1.0e-13, np.sign(ct), ct)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = np.where(np.abs(x) < 1.0e-10, 0.5 * math.pi * np.where(y >= 0.0, 1.0, -1.0),
                       np.arctan(y / np.where(x == 0.0, 1.0, x)))
    phi = np.where((np.abs(x) >= 1.0e-10) & (x < 0.0), phi + math.pi, phi)
    # P_L^M(ct) by the upward recurrence
    pmm = np.ones_like(ct)
    if M > 0:
        somx2 = np.sqrt((1.0 - ct) * (1.0 + ct))
        fact = 1.0
        for _ in range(M):
            pmm = -pmm * fact * somx2
            fact += 2.0
    if L == M:
        p = pmm
    else:
        pmmp1 = ct * (2.0 * M + 1.0) * pmm
        if L == M + 1:
            p = pmmp1
        else:
            for ll in range(M + 2, L + 1):
                pll = (ct * (2.0 * ll - 1.0) * pmmp1 - (ll + M - 1.0) * pmm) / (ll - M)
                pmm, pmmp1 = pmmp1, pll
            p = pll
    if M_in > 0:
        v = math.sqrt(2.0) * clm * p * np.cos(phi * M)
    elif M_in == 0:
        v = clm * p
    elif M % 2 == 0:
        v = -math.sqrt(2.0) * clm * p * np.sin(phi * M)
    else:
        v = math.sqrt(2.0) * clm * p * np.sin(phi * M)
    return np.where(origin, 0.0, v)


def radial_at(fp: FilteredProjectors, k: int, r):
    """Projector k at distances r (prep_ps_nloc2's interpolation,
    ps_nloc2_module.f90:306-339): bisection for the mesh interval, the value
    of point 2 below r_3, else the 20-window Neville search stopping once the
    error estimate is below 1e-8."""
    r = np.asarray(r, dtype=np.float64)
    nrc = int(fp.nrps[k])
    rad1 = fp.rad1[:nrc]
    f = fp.viod[:nrc, k]
    # ir: largest 1-based index with rad1(ir) <= r among 1..nrc-1 (the bisection)
    ir = np.searchsorted(rad1, r, side="right")
    ir = np.clip(ir, 1, nrc - 1)
    out = np.empty(len(r))
    small = ir <= 2
    out[small] = f[1]
    big = ~small
    if np.any(big):
        v, _ = _windowed_interp(rad1, f, r[big], ir[big] - 1, 0, nrc - 1, ep=1.0e-8)
        out[big] = v
    return out


def projector_tables(grid: Grid, frac: np.ndarray, fp: FilteredProjectors,
                     cell: np.ndarray) -> Projectors:
    """KB tables of every atom (positions ``frac`` in units of the cell
    vectors, rows of ``cell``): the grid points within the cut-off of the
    atom (periodic images folded back onto the grid, values of one point
    summed, zeros dropped, points sorted by grid index), uV = v_l(r) Y_lm,
    lma ordered atom, projector, m = -l..l (ps_nloc2_module.f90:197-345,
    :569-607, :1128-1190)."""
    N = np.array(grid.N)
    aa = np.asarray(cell, dtype=np.float64).T          # columns = cell vectors
    h = np.array(grid.h)
    rmax = float(np.max(fp.rps))
    M = [int(math.ceil(rmax / h[d])) + 2 for d in range(3)]
    o3, o2, o1 = np.meshgrid(np.arange(-M[2], M[2] + 1), np.arange(-M[1], M[1] + 1),
                             np.arange(-M[0], M[0] + 1), indexing="ij")
    offs = np.stack([o1.ravel(), o2.ravel(), o3.ravel()], axis=1)
    lists = []
    amap, lmap, mmap, iuv = [], [], [], []
    full = []
    for a in range(len(frac)):
        R = aa @ frac[a]
        ic = np.floor(frac[a] * N + 0.5).astype(np.int64)     # nint, positive coordinates
        ic = np.where(frac[a] * N < 0, -np.floor(-frac[a] * N + 0.5), ic).astype(np.int64)
        idg = ic[None, :] + offs
        d = idg * (1.0 / N)[None, :]
        xyz = d @ aa.T - R[None, :]
        r2 = np.sum(xyz * xyz, axis=1)
        wrapped = np.mod(idg, N[None, :])
        gidx = wrapped[:, 0] + N[0] * (wrapped[:, 1] + N[1] * wrapped[:, 2])
        for k in range(len(fp.lo)):
            L = int(fp.lo[k])
            keep = r2 <= fp.rps[k] ** 2 + 1.0e-10
            x = xyz[keep]
            r = np.sqrt(r2[keep])
            at_center = ((np.abs(x[:, 0]) <= 1.0e-14) & (np.abs(x[:, 1]) <= 1.0e-14)
                         & (np.abs(x[:, 2]) <= 1.0e-14))
            v0 = np.zeros(len(r))
            interp = ~at_center if L > 0 else np.ones(len(r), bool)
            if np.any(interp):
                v0[interp] = radial_at(fp, k, r[interp])
            gi = gidx[keep]
            for m in range(-L, L + 1):
                uv = v0 * real_ylm(x[:, 0], x[:, 1], x[:, 2], L, m)
                # merge images of one point in encounter order: a point enters
                # with its first nonzero value, later copies add to it
                order = np.argsort(gi, kind="stable")
                si, sv = gi[order], uv[order]
                full.append(si)
                uniq, start = np.unique(si, return_index=True)
                summed = np.add.reduceat(sv, start) if len(sv) else sv
                nonzero = np.logical_or.reduceat(sv != 0.0, start) if len(sv) else sv
                lists.append((uniq[nonzero] + 1, summed[nonzero]))
                amap.append(a + 1)
                lmap.append(L)
                mmap.append(m)
                iuv.append(int(fp.inorm[k]))
    nz = len(lists)
    maxmjj = max(len(li[0]) for li in lists)
    mmjj = max(len(f) for f in full)
    MJJ = np.zeros(nz, np.int32)
    JJP = np.zeros((nz, maxmjj), np.int32)
    uVk = np.zeros((nz, maxmjj))
    MJJ_MAP = np.zeros(nz, np.int32)
    JJ_MAP = np.zeros((nz, mmjj), np.int32)
    for k, (ji, jv) in enumerate(lists):
        MJJ[k] = len(ji)
        JJP[k, :len(ji)] = ji
        uVk[k, :len(ji)] = jv
    for k, fi in enumerate(full):
        MJJ_MAP[k] = len(fi)
        JJ_MAP[k, :len(fi)] = fi + 1
    return Projectors(nz, MJJ, JJP, uVk, np.array(iuv, np.int32), np.array(amap, np.int32),
                      np.array(lmap, np.int32), np.array(mmap, np.int32), MJJ_MAP, JJ_MAP,
                      np.zeros((nz, mmjj, 3)), float("nan"))


# ----------------------------------------------------------------------------
# G sphere, structure factor, local potential, initial density
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class GSphere:
    """The reference's G sphere (ggrid_module.f90:55-150, orthorhombic cell):
    integer vectors |i_d| <= (N_d-1)/2 with G^2 < Ecut - 1e-14, grouped into
    shells of equal |G|^2 (consecutive sorted values within 1e-12)."""
    ecut: float
    gcut: float
    idx: np.ndarray            # (NG, 3) integer components
    g2: np.ndarray             # (NG,)
    shell: np.ndarray          # (NG,) shell of each vector
    gg: np.ndarray             # (NMGL,) |G|^2 of each shell


def g_sphere(grid: Grid) -> GSphere:
    b = np.array([2.0 * math.pi / L for L in grid.L])
    ng = np.array([(n - 1) // 2 for n in grid.N])
    gcut = 0.5 * float(np.min(b * (2 * ng + 1)))
    ecut = gcut * gcut
    i3, i2, i1 = np.meshgrid(np.arange(-ng[2], ng[2] + 1), np.arange(-ng[1], ng[1] + 1),
                             np.arange(-ng[0], ng[0] + 1), indexing="ij")
    idx = np.stack([i1.ravel(), i2.ravel(), i3.ravel()], axis=1)
    g2 = np.sum((idx * b[None, :]) ** 2, axis=1)
    keep = g2 < ecut - 1.0e-14
    idx, g2 = idx[keep], g2[keep]
    order = np.argsort(g2, kind="stable")
    sg = g2[order]
    new = np.concatenate([[False], np.abs(np.diff(sg)) > 1.0e-12])
    new[0] = sg[0] > 1.0e-12 if len(sg) else False
    sh_sorted = np.cumsum(new)
    shell = np.empty(len(g2), np.int64)
    shell[order] = sh_sorted
    nshell = int(sh_sorted[-1]) + 1
    gg = np.zeros(nshell)
    gg[sh_sorted] = sg                       # the last member of each run (as GG(n)=s2)
    return GSphere(ecut, gcut, idx, g2, shell, gg)


def structure_factor(gs: GSphere, frac: np.ndarray) -> np.ndarray:
    """S(G) = sum_a exp(-i 2 pi G.tau_a) (strfac_module.f90:15-37)."""
    phase = 2.0 * math.pi * (gs.idx.astype(np.float64) @ frac.T)     # (NG, natom)
    return np.sum(np.cos(phase), axis=1) - 1j * np.sum(np.sin(phase), axis=1)


def to_real_space(grid: Grid, gs: GSphere, coef: np.ndarray) -> np.ndarray:
    """Re sum_G coef(G) exp(+i G.r) on the grid (the unnormalised backward
    transform of ps_local_fftw_module.f90:34-77), returned x-fastest (ML,)."""
    N1, N2, N3 = grid.N
    z = np.zeros((N3, N2, N1), np.complex128)
    z[gs.idx[:, 2] % N3, gs.idx[:, 1] % N2, gs.idx[:, 0] % N1] = coef
    return np.real(np.fft.ifftn(z)).ravel() * (N1 * N2 * N3)


def _sinc_series(x):
    """sin(x)/x as the reference evaluates it (power series below 0.1)."""
    small = x < 0.1
    out = np.empty_like(x)
    xs = x[small]
    out[small] = -(1.0 / 39916800.0 * xs ** 10 - 1.0 / 362880.0 * xs ** 8 + 1.0 / 5040.0 * xs ** 6
                   - 1.0 / 120.0 * xs ** 4 + 1.0 / 6.0 * xs ** 2 - 1.0)
    xb = x[~small]
    out[~small] = np.sin(xb) / xb
    return out


def local_form_factor(ps: Pseudopotential, gg: np.ndarray, vcell: float, parloc=None):
    """v_loc(|G|) per shell (init_ps_local, ps_local_module.f90:180-300):
    (4 pi / V) [int r^2 (v - v_long) j0(G r) dr (Simpson on the log mesh)
    + the transform of v_long = -Z/r [p1 erf(sqrt(a2) r) + p3 erf(sqrt(a4) r)]
    in closed form].  The split (p1, a2, p3, a4) defaults to the file's; the
    reference refits it (simc, MINPACK lmder1 on samples of v), but the sum
    does not depend on it beyond the quadrature of the short-range remainder
    (1.7e-14 Ha between the file's split and a refit one for Si)."""
    from scipy.special import erf
    if parloc is None:
        parloc = ps.parloc
    z = ps.zps
    p1, p2, p3, p4 = parloc[0], math.sqrt(parloc[1]), parloc[2], math.sqrt(parloc[3])
    r = ps.rad
    with np.errstate(divide="ignore", invalid="ignore"):
        vlong = np.where(r < 1.0e-9, -2.0 * z / math.sqrt(math.pi) * (p1 * p2 + p3 * p4),
                         -z / np.where(r == 0.0, 1.0, r) * (p1 * erf(p2 * r) + p3 * erf(p4 * r)))
    vshort = ps.vql - vlong
    const = FOUR_PI / vcell
    out = np.zeros(len(gg))
    base = r * r * vshort * ps.rab
    for ig, g2 in enumerate(gg):
        G = math.sqrt(g2)
        f = base if G == 0.0 else base * _sinc_series(G * r)
        out[ig] = simpson(f) * const
    q1, q2 = -z * parloc[0], 0.25 / parloc[1]
    q3, q4 = -z * parloc[2], 0.25 / parloc[3]
    with np.errstate(divide="ignore", invalid="ignore"):
        longg = np.where(gg == 0.0, -(q1 * q2 + q3 * q4),
                         (q1 * np.exp(-gg * q2) + q3 * np.exp(-gg * q4)) / np.where(gg == 0, 1, gg))
    return out + longg * const


def initrho_form_factor(ps: Pseudopotential, gg: np.ndarray, vcell: float):
    """(1/V) int cdd(r) j0(G r) dr per shell (init_ps_initrho_g,
    ps_initrho_module.f90:96-144)."""
    out = np.zeros(len(gg))
    base = ps.cdd * ps.rab
    for ig, g2 in enumerate(gg):
        G = math.sqrt(g2)
        f = base if G == 0.0 else base * _sinc_series(G * ps.rad)
        out[ig] = simpson(f) / vcell
    return out


# ----------------------------------------------------------------------------
# Ewald energy (ewald_module.f90:604-820)
# ----------------------------------------------------------------------------

def ewald_energy(cell: np.ndarray, frac: np.ndarray, z: np.ndarray, eta: float | None = None,
                 tol: float = 1.0e-17) -> float:
    """Ion-ion energy of point charges z at frac (rows of ``cell`` are the
    lattice vectors) in a neutralising background: Ewald's split with
    parameter eta (default pi / V^(2/3) as the reference), real-space
    sum_{pairs, images} z_i z_j erfc(sqrt(eta) r) / r over images while erfc
    exceeds ``tol``, reciprocal sum (4 pi / V) sum_G |S(G)|^2 exp(-G^2/4eta)/G^2
    while the Gaussian exceeds ``tol``, minus the self term 2 sqrt(eta/pi)
    sum z^2 and the background term pi Q^2 / (eta V); halved."""
    from scipy.special import erfc
    A = np.asarray(cell, dtype=np.float64)
    V = abs(float(np.linalg.det(A)))
    if eta is None:
        eta = math.pi / V ** (2.0 / 3.0)
    s = math.sqrt(eta)
    pos = np.asarray(frac) @ A
    z = np.asarray(z, dtype=np.float64)
    # real space: images n with min distance below rcut
    x = 1.0
    while math.erfc(s * x) > tol:
        x *= 1.1
    rcut = x
    B = 2.0 * math.pi * np.linalg.inv(A).T
    nmax = [int(math.ceil(rcut * np.linalg.norm(B[d]) / (2.0 * math.pi))) + 1 for d in range(3)]
    er = 0.0
    d0 = pos[:, None, :] - pos[None, :, :]
    zz = z[:, None] * z[None, :]
    for n1 in range(-nmax[0], nmax[0] + 1):
        for n2 in range(-nmax[1], nmax[1] + 1):
            for n3 in range(-nmax[2], nmax[2] + 1):
                t = n1 * A[0] + n2 * A[1] + n3 * A[2]
                d = np.sqrt(np.sum((d0 + t[None, None, :]) ** 2, axis=2))
                if n1 == 0 and n2 == 0 and n3 == 0:
                    np.fill_diagonal(d, np.inf)
                near = d < rcut
                er += float(np.sum(zz[near] * erfc(s * d[near]) / d[near]))
    # reciprocal space
    g = 1.0
    while math.exp(-g * g / (4.0 * eta)) > tol:
        g *= 1.1
    gcut = g
    mmax = [int(math.ceil(gcut / np.linalg.norm(B[d]))) + 1 for d in range(3)]
    m3, m2, m1 = np.meshgrid(*(np.arange(-mmax[d], mmax[d] + 1) for d in (2, 1, 0)),
                             indexing="ij")
    mv = np.stack([m1.ravel(), m2.ravel(), m3.ravel()], axis=1)
    mv = mv[np.any(mv != 0, axis=1)]
    G = mv @ B
    g2 = np.sum(G * G, axis=1)
    sel = g2 <= gcut * gcut
    G, g2, mv = G[sel], g2[sel], mv[sel]
    phase = 2.0 * math.pi * (mv.astype(np.float64) @ np.asarray(frac).T)
    S2 = (np.cos(phase) @ z) ** 2 + (np.sin(phase) @ z) ** 2
    eg = FOUR_PI / V * float(np.sum(S2 * np.exp(-g2 / (4.0 * eta)) / g2))
    q, q2 = float(np.sum(z)), float(np.sum(z * z))
    eg -= math.pi / (eta * V) * q * q + 2.0 * math.sqrt(eta / math.pi) * q2
    return 0.5 * (er + eg)


# ----------------------------------------------------------------------------
# the whole set-up
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class PsvSystem:
    system: System             # grid, stencil, projector tables, Vion as vloc, occupations
    vion: np.ndarray           # (ML,) ionic local potential
    rho0: np.ndarray           # (ML,) initial density, normalised to nelectron
    rho0_raw: tuple            # (sum*dV, min, max) before normalisation
    nelectron: float
    eewald: float
    gsphere: GSphere
    filtered: FilteredProjectors
    parloc: np.ndarray


def make_psv_system(ps: Pseudopotential, cell: np.ndarray, frac: np.ndarray, N,
                    Md: int = 6, nband: int | None = None, seed: int = 20240601,
                    rcfac: float = 1.5, qcfac: float = 1.0, etafac: float = 8.0,
                    name: str = "psv") -> PsvSystem:
    """A one-element periodic system with a PSV pseudopotential, set up as the
    reference does before its first sweep (init_ps_local, construct_strfac,
    construct_ps_local, construct_ps_initrho, ps_nloc_initiate, init_eion);
    Nband defaults to the reference's choice for an insulator (1.5 x the
    occupied bands, check_Nband_electron)."""
    cell = np.asarray(cell, dtype=np.float64)
    if np.any(np.abs(cell - np.diag(np.diag(cell))) > 0.0):
        raise ValueError("the stencil path handles orthorhombic cells only")
    frac = np.asarray(frac, dtype=np.float64)
    L = tuple(float(cell[d, d]) for d in range(3))
    grid = Grid(tuple(int(n) for n in N), L)
    vcell = grid.ML * grid.dV
    gs = g_sphere(grid)
    fp = filter_projectors(ps, gs.gcut, rcfac, qcfac, etafac)
    proj = projector_tables(grid, frac, fp, cell)
    sgk = structure_factor(gs, frac)
    parloc = np.array(ps.parloc)
    vqlg = local_form_factor(ps, gs.gg, vcell, parloc)
    vion = to_real_space(grid, gs, vqlg[gs.shell] * sgk)
    cddg = initrho_form_factor(ps, gs.gg, vcell)
    rho = to_real_space(grid, gs, cddg[gs.shell] * sgk)
    rho = np.where(rho < 0.0, 0.0, rho)
    raw = (float(np.sum(rho) * grid.dV), float(np.min(rho)), float(np.max(rho)))
    natom = len(frac)
    nel = ps.zps * natom
    rho = rho * (nel / raw[0])
    eew = ewald_energy(cell, frac, np.full(natom, float(round(ps.zps))))
    nocc = int(round(nel / 2.0))
    if nband is None:
        nband = int(nel / 2.0 * 1.5)
    occ = np.zeros(nband)
    occ[:nocc] = 2.0
    h = np.array(grid.h)
    pos = frac @ cell
    ic = np.rint(pos / h[None, :]).astype(np.int64)
    delta = pos - ic * h[None, :]
    ic = np.mod(ic, np.array(grid.N)[None, :])
    c0, cl = kinetic_coefficients(Md, grid.h)
    system = System(name, grid, Md, ic, delta, np.zeros((natom, 4)), vion, proj, nband, occ,
                    seed, c0, cl)
    return PsvSystem(system, vion, rho, raw, nel, eew, gs, fp, parloc)
