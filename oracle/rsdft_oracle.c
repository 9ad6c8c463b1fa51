/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64 restatement of RSDFT's Kohn-Sham inner loop at fixed
 * potential (Gamma point, real wavefunctions, one MPI rank).  It is the
 * checker for the HIP library and the CPU baseline timed beside it; nothing
 * in the product links or calls it.  Each function names the reference
 * lines (paths relative to the RSDFT source tree) whose algorithm it follows.
 *
 * Layout: psi is band-major, psi[b*ML + i] with i = i1 + N1*(i2 + N2*i3);
 * coef_lap is Fortran coef_lap(3,Md): element (d,m) at [d + 3*(m-1)];
 * projector tables are Fortran JJP/uVk(MAXMJJ,nzlma): (j,lma) at
 * [lma*MAXMJJ + j], JJP 1-based.
 *
 * Parity pins (tests/test_oracle_*.py): FD weights against the closed form
 * and the reference's printed table; kinetic operator against exact
 * plane-wave eigenvalues and a brute-force dense matrix; nonlocal against a
 * dense sum D*beta*beta^T*dV; full H against Hermiticity and dense
 * diagonalisation; Gram-Schmidt against Cholesky-QR; the preconditioner
 * against an explicit dense CG; the 2x2 solver against the generalized
 * eigen-decomposition from scipy; CG against monotone Rayleigh quotients and
 * convergence to dense eigenpairs; density and force against central finite
 * differences of the band energy.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int N1, N2, N3, Md;
    double h[3];
    double dV;
    double coef_lap0;
    const double *coef_lap;   /* (3, Md) Fortran order */
    const double *vloc;       /* (ML) */
    int nzlma, maxmjj;
    const int *MJJ;           /* (nzlma) */
    const int *JJP;           /* (MAXMJJ, nzlma), 1-based */
    const double *uVk;        /* (MAXMJJ, nzlma) */
    const int *iuV;           /* (nzlma) */
} orc_ham;

static inline int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }

/* get_coef_lapla_fd (fd_module.f90:18-54): lap(-Md..Md) from Lagrange
 * products; lap is an array of 2*Md+1 doubles, lap[Md + j] = lap(j). */
void orc_fd_coef(int Md, double *lap)
{
    double s = 0.0;
    for (int k = 0; k < 2 * Md + 1; ++k) lap[k] = 0.0;
    for (int i = -Md; i <= Md; ++i) {
        if (i == 0) continue;
        for (int j = -Md; j <= Md; ++j) {
            if (j == i || j == 0) continue;
            s += 1.0 / (double)(i * j);
        }
    }
    lap[Md] = s;
    for (int j = 1; j <= Md; ++j) {
        double t = 1.0;
        for (int i = -Md; i <= Md; ++i) {
            if (i == j) continue;
            t *= (double)(j - i);
        }
        s = 0.0;
        for (int k = -Md; k <= Md; ++k) {
            if (k == j || k == 0) continue;
            double s0 = 1.0;
            for (int i = -Md; i <= Md; ++i) {
                if (i == k || i == j || i == 0) continue;
                s0 *= (double)(-i);
            }
            s += s0;
        }
        lap[Md + j] = 2.0 * s / t;
        lap[Md - j] = lap[Md + j];
    }
}

/* op_kinetic_sol, Gamma / orthorhombic branch (kinetic_sol_module.f90:72-77,
 * 129-149): htpsi is OVERWRITTEN with coef_lap0*tpsi, then the m-outer
 * accumulation of the three axis terms; the halo of www is the periodic
 * wrap of bc_module.f90:722-736. */
void orc_op_kinetic(const orc_ham *H, int nb, const double *tpsi, double *htpsi)
{
    const int N1 = H->N1, N2 = H->N2, N3 = H->N3, Md = H->Md;
    const long ML = (long)N1 * N2 * N3;
    for (int ib = 0; ib < nb; ++ib) {
        const double *f = tpsi + ib * ML;
        double *hf = htpsi + ib * ML;
        const double c = H->coef_lap0;   /* + const_k2(k) = 0 at Gamma */
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < ML; ++i) hf[i] = c * f[i];
        for (int m = 1; m <= Md; ++m) {
            const double c1 = H->coef_lap[0 + 3 * (m - 1)];
            const double c2 = H->coef_lap[1 + 3 * (m - 1)];
            const double c3 = H->coef_lap[2 + 3 * (m - 1)];
            #pragma omp parallel for schedule(static)
            for (int i3 = 0; i3 < N3; ++i3) {
                const int z0 = wrap(i3 - m, N3), z1 = wrap(i3 + m, N3);
                for (int i2 = 0; i2 < N2; ++i2) {
                    const int y0 = wrap(i2 - m, N2), y1 = wrap(i2 + m, N2);
                    for (int i1 = 0; i1 < N1; ++i1) {
                        const int x0 = wrap(i1 - m, N1), x1 = wrap(i1 + m, N1);
                        const long j = i1 + (long)N1 * (i2 + (long)N2 * i3);
                        hf[j] = hf[j]
                              + c1 * (f[x0 + (long)N1 * (i2 + (long)N2 * i3)] + f[x1 + (long)N1 * (i2 + (long)N2 * i3)])
                              + c2 * (f[i1 + (long)N1 * (y0 + (long)N2 * i3)] + f[i1 + (long)N1 * (y1 + (long)N2 * i3)])
                              + c3 * (f[i1 + (long)N1 * (i2 + (long)N2 * z0)] + f[i1 + (long)N1 * (i2 + (long)N2 * z1)]);
                    }
                }
            }
        }
    }
}

/* op_localpot (localpot_module.f90:70-76): vf += Vloc*f. */
void orc_op_localpot(const orc_ham *H, int nb, const double *f, double *vf)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    for (int n = 0; n < nb; ++n) {
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < ML; ++i) vf[n * ML + i] += H->vloc[i] * f[n * ML + i];
    }
}

/* op_ps_nloc2_hp on one rank (ps_nloc2_op_module.f90:161-175, 255-263):
 * uVunk(lma) = iuV*dV*sum_j uVk(j,lma)*tpsi(JJP(j,lma)), then
 * htpsi(JJP(j,lma)) += uVk(j,lma)*uVunk(lma).  Accumulates. */
void orc_op_nonlocal(const orc_ham *H, int nb, const double *tpsi, double *htpsi)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    const int nz = H->nzlma, mx = H->maxmjj;
    if (nz <= 0) return;
    double *uVunk = (double *)malloc(sizeof(double) * nz);
    for (int ib = 0; ib < nb; ++ib) {
        const double *f = tpsi + ib * ML;
        double *hf = htpsi + ib * ML;
        #pragma omp parallel for schedule(static)
        for (int lma = 0; lma < nz; ++lma) {
            double t = 0.0;
            for (int j = 0; j < H->MJJ[lma]; ++j)
                t += H->uVk[(long)lma * mx + j] * f[H->JJP[(long)lma * mx + j] - 1];
            uVunk[lma] = H->iuV[lma] * H->dV * t;
        }
        for (int lma = 0; lma < nz; ++lma)
            for (int j = 0; j < H->MJJ[lma]; ++j)
                hf[H->JJP[(long)lma * mx + j] - 1] += H->uVk[(long)lma * mx + j] * uVunk[lma];
    }
    free(uVunk);
}

/* hamiltonian (hamiltonian_module.f90:20-72): zero, kinetic (overwrite),
 * local, nonlocal; the Fock term is zero. */
void orc_hamiltonian(const orc_ham *H, int nb, const double *tpsi, double *htpsi)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    memset(htpsi, 0, sizeof(double) * ML * nb);
    orc_op_kinetic(H, nb, tpsi, htpsi);
    orc_op_localpot(H, nb, tpsi, htpsi);
    orc_op_nonlocal(H, nb, tpsi, htpsi);
}

static double dot(long n, const double *a, const double *b)
{
    double t = 0.0;
    #pragma omp parallel for reduction(+:t) schedule(static)
    for (long i = 0; i < n; ++i) t += a[i] * b[i];
    return t;
}

/* precond_cg_mat2_omp with E = 0 (cgpc_hprsdft_module.f90:281-348):
 * f = c*g + sum_d c_d (g(+e_d) + g(-e_d)), c = sum 1/h^2, c_d = -0.5/h_d^2. */
void orc_precond_apply(const orc_ham *H, const double *g, double *f)
{
    const int N1 = H->N1, N2 = H->N2, N3 = H->N3;
    const double c = 1.0 / (H->h[0] * H->h[0]) + 1.0 / (H->h[1] * H->h[1]) + 1.0 / (H->h[2] * H->h[2]);
    const double c1 = -0.5 / (H->h[0] * H->h[0]), c2 = -0.5 / (H->h[1] * H->h[1]),
                 c3 = -0.5 / (H->h[2] * H->h[2]);
    const double d = c - 0.0;
    #pragma omp parallel for schedule(static)
    for (int i3 = 0; i3 < N3; ++i3)
        for (int i2 = 0; i2 < N2; ++i2)
            for (int i1 = 0; i1 < N1; ++i1) {
                const long i = i1 + (long)N1 * (i2 + (long)N2 * i3);
                double t = d * g[i];
                t = t + c1 * (g[wrap(i1 + 1, N1) + (long)N1 * (i2 + (long)N2 * i3)] + g[wrap(i1 - 1, N1) + (long)N1 * (i2 + (long)N2 * i3)])
                      + c2 * (g[i1 + (long)N1 * (wrap(i2 + 1, N2) + (long)N2 * i3)] + g[i1 + (long)N1 * (wrap(i2 - 1, N2) + (long)N2 * i3)])
                      + c3 * (g[i1 + (long)N1 * (i2 + (long)N2 * wrap(i3 + 1, N3))] + g[i1 + (long)N1 * (i2 + (long)N2 * wrap(i3 - 1, N3))]);
                f[i] = t;
            }
}

/* precond_hprsdft for one band (cgpc_hprsdft_module.f90:126-263): mloop
 * CG steps on A x = g from x0 = g (Pgk holds g on entry); work is 3*ML. */
void orc_precond(const orc_ham *H, int mloop, const double *gk, double *Pgk, double *work)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    double *ftmp = work, *rk = work + ML, *pk = work + 2 * ML;
    if (mloop <= 0) return;
    orc_precond_apply(H, gk, ftmp);
    for (long i = 0; i < ML; ++i) { rk[i] = gk[i] - ftmp[i]; pk[i] = rk[i]; }
    double rr1 = dot(ML, rk, rk);
    if (rr1 < 1.0e-24) return;
    for (int iloop = 1; iloop <= mloop; ++iloop) {
        orc_precond_apply(H, pk, ftmp);
        const double pAp = dot(ML, pk, ftmp);
        const double rr0 = rr1;
        const double a = rr0 / pAp;
        for (long i = 0; i < ML; ++i) Pgk[i] = Pgk[i] + a * pk[i];
        if (iloop == mloop) break;
        for (long i = 0; i < ML; ++i) rk[i] = rk[i] - a * ftmp[i];
        rr1 = dot(ML, rk, rk);
        const double b = rr1 / rr0;
        for (long i = 0; i < ML; ++i) pk[i] = rk[i] + b * pk[i];
    }
}

/* The 2x2 problem dsygv(1,'V','U') solves in conjugate_gradient_1
 * (cg_module.f90:318): A z = w B z, w ascending, z^T B z = 1, computed as
 * LAPACK does it -- Cholesky B = U^T U, C = U^-T A U^-1, symmetric eigen of
 * C by one Jacobi rotation, z = U^-1 y.  A, B, Z are column-major 2x2. */
void orc_sygv2(const double *A, const double *B, double *W, double *Z)
{
    const double u11 = sqrt(B[0]);
    const double u12 = B[2] / u11;
    const double u22 = sqrt(B[3] - u12 * u12);
    /* Ui = U^-1 = [[1/u11, -u12/(u11 u22)], [0, 1/u22]] */
    const double i11 = 1.0 / u11, i12 = -u12 / (u11 * u22), i22 = 1.0 / u22;
    const double a = A[0], b = A[2], c = A[3];
    /* C = Ui^T A Ui */
    const double c11 = i11 * a * i11;
    const double c12 = i11 * (a * i12 + b * i22);
    const double c22 = i12 * (a * i12 + b * i22) + i22 * (b * i12 + c * i22);
    double w1, w2, y11, y21, y12, y22;
    if (c12 == 0.0) {
        w1 = c11; w2 = c22; y11 = 1.0; y21 = 0.0; y12 = 0.0; y22 = 1.0;
    } else {
        const double tau = (c22 - c11) / (2.0 * c12);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        w1 = c11 - t * c12; w2 = c22 + t * c12;
        y11 = cs; y21 = -sn; y12 = sn; y22 = cs;
    }
    if (w2 < w1) {
        double tw = w1; w1 = w2; w2 = tw;
        double t1 = y11, t2 = y21; y11 = y12; y21 = y22; y12 = t1; y22 = t2;
    }
    W[0] = w1; W[1] = w2;
    Z[0] = i11 * y11 + i12 * y21; Z[1] = i22 * y21;
    Z[2] = i11 * y12 + i12 * y22; Z[3] = i22 * y22;
}

/* conjugate_gradient_1, real version (cg_module.f90:97-430), one band per
 * step (MB_d = 1), SWGS = 0 (cggs is a no-op), preconditioner IPC = 4.
 * unk is updated in place; esp and res are written.  work: 9*ML doubles.
 * counts (optional, 2 ints): hamiltonian calls and rejected updates. */
void orc_cg(const orc_ham *H, int MB, int Ncg, int mloop, double *unk, double *esp,
            double *res, double *work, int *counts)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    const double dV = H->dV, c1 = 2.0;
    double *hxk = work, *hpk = work + ML, *gk = work + 2 * ML, *Pgk = work + 3 * ML,
           *pk = work + 4 * ML, *pcw = work + 5 * ML;
    int nh = 0, nrej = 0;
    for (int n = 0; n < MB; ++n) { res[n] = 0.0; esp[n] = 0.0; }
    for (int ns = 0; ns < MB; ++ns) {
        double *x = unk + ns * ML;
        double E1 = 1.0e10, E, rb, gkgk = 0.0;
        orc_hamiltonian(H, 1, x, hxk); ++nh;
        E = dot(ML, x, hxk) * dV;
        for (long i = 0; i < ML; ++i) gk[i] = -c1 * (hxk[i] - E * x[i]);
        rb = dot(ML, gk, gk) * dV;
        for (int icg = 1; icg <= Ncg + 1; ++icg) {
            memcpy(Pgk, gk, sizeof(double) * ML);
            res[ns] = rb / (c1 * c1);
            if (rb < 0.0) break;
            if (fabs(E - E1) < 1.0e-15) break;
            if (icg == Ncg + 1) break;
            orc_precond(H, mloop, gk, Pgk, pcw);
            rb = dot(ML, Pgk, gk) * dV;
            if (icg == 1) {
                memcpy(pk, Pgk, sizeof(double) * ML);
            } else {
                const double bk = rb / gkgk;
                for (long i = 0; i < ML; ++i) pk[i] = Pgk[i] + bk * pk[i];
            }
            gkgk = rb;
            orc_hamiltonian(H, 1, pk, hpk); ++nh;
            double v[6];
            v[0] = dot(ML, x, x) * dV;
            v[1] = dot(ML, pk, x) * dV;
            v[2] = dot(ML, pk, pk) * dV;
            v[3] = dot(ML, x, hxk) * dV;
            v[4] = dot(ML, pk, hxk) * dV;
            v[5] = dot(ML, pk, hpk) * dV;
            const double Bm[4] = {v[0], v[1], v[1], v[2]};
            const double Am[4] = {v[3], v[4], v[4], v[5]};
            double W[2], Z[4];
            orc_sygv2(Am, Bm, W, Z);
            if (fabs(W[0] - E) > 1.0e-1 && fabs(W[1] - E) <= 1.0e-1) {
                Z[0] = Z[2]; Z[1] = Z[3]; W[0] = W[1];
            }
            if (Z[0] < 0.0) { Z[0] = -Z[0]; Z[1] = -Z[1]; }
            const double u1 = Z[0], u2 = Z[1], w = W[0];
            E1 = E;
            E = w;
            for (long i = 0; i < ML; ++i) hxk[i] = u1 * hxk[i] + u2 * hpk[i];
            for (long i = 0; i < ML; ++i) gk[i] = -c1 * (hxk[i] - w * (u1 * x[i] + u2 * pk[i]));
            rb = dot(ML, gk, gk) * dV;
            if (rb / res[ns] > 1.0e8) { E = E1; ++nrej; continue; }
            for (long i = 0; i < ML; ++i) x[i] = u1 * x[i] + u2 * pk[i];
        }
        esp[ns] = E;
    }
    if (counts) { counts[0] = nh; counts[1] = nrej; }
}

/* calc_density, NCPP branch (density_module.f90:129-136): rho = sum occ*psi^2. */
void orc_density(long ML, int MB, const double *occ, const double *unk, double *rho)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < ML; ++i) {
        double t = 0.0;
        for (int n = 0; n < MB; ++n) t += occ[n] * unk[n * ML + i] * unk[n * ML + i];
        rho[i] = t;
    }
}

/* calc_force_ps_nloc2, real psi on one rank (ps_nloc2_module.f90:1601-1682,
 * 1760-1773): per occupied band, w0 = iuV*(-2 occ dV^2)*sum uVk psi(JJP),
 * w = sum_{j<MJJ_MAP} duVdR(:,j) psi(JJ_MAP(j)), force(:,a) += w0*w.
 * duVdR(3, mmjj, nzlma) Fortran order, JJ_MAP(mmjj, nzlma) 1-based. */
void orc_force_nloc(const orc_ham *H, int MB, int natom, const double *occ, const double *unk,
                    const int *amap, const int *MJJ_MAP, int mmjj, const int *JJ_MAP,
                    const double *duVdR, double *force)
{
    const long ML = (long)H->N1 * H->N2 * H->N3;
    const int nz = H->nzlma, mx = H->maxmjj;
    for (int k = 0; k < 3 * natom; ++k) force[k] = 0.0;
    for (int n = 0; n < MB; ++n) {
        if (occ[n] == 0.0) continue;
        const double c = -2.0 * occ[n] * H->dV * H->dV;
        const double *f = unk + n * ML;
        for (int lma = 0; lma < nz; ++lma) {
            double w0 = 0.0, w1 = 0.0, w2 = 0.0, w3 = 0.0;
            for (int j = 0; j < H->MJJ[lma]; ++j)
                w0 += H->uVk[(long)lma * mx + j] * f[H->JJP[(long)lma * mx + j] - 1];
            w0 = H->iuV[lma] * c * w0;
            for (int j = 0; j < MJJ_MAP[lma]; ++j) {
                const double p = f[JJ_MAP[(long)lma * mmjj + j] - 1];
                const double *d = duVdR + 3 * ((long)lma * mmjj + j);
                w1 += d[0] * p; w2 += d[1] * p; w3 += d[2] * p;
            }
            const int a = amap[lma] - 1;
            force[3 * a + 0] += w0 * w1;
            force[3 * a + 1] += w0 * w2;
            force[3 * a + 2] += w0 * w3;
        }
    }
}
