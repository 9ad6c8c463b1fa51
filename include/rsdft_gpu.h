/*
 * rsdft_gpu.h -- C-ABI of the MI355X Kohn-Sham inner-loop library.
 *
 * Every entry point returns 0 on success and a non-zero status otherwise
 * (rsdft_gpu_last_error() gives the message).  Arrays are plain host
 * pointers in the reference's Fortran layout:
 *   psi(n1:n2, ib1:ib2)   band-major, grid index i1 + N1*(i2 + N2*i3);
 *   coef_lap(3, Md)       Fortran column-major;
 *   JJP/uVk(MAXMJJ,nzlma) Fortran column-major, JJP 1-based;
 *   duVdR(3,MMJJ,nzlma)   Fortran column-major.
 * One process drives one GPU; the library is not re-entrant.  All work is
 * ordered on one HIP stream.
 *
 * Residency: rsdft_gpu_wf_attach() gives the library a device mirror of a
 * host wavefunction array.  The operations below that take `unk` operate on
 * that mirror when `unk` is the attached pointer (no host traffic; the host
 * copy is refreshed only by rsdft_gpu_wf_sync(0)); any other pointer is
 * copied in and out around the call.
 */
#ifndef RSDFT_GPU_H
#define RSDFT_GPU_H

#if defined(__GNUC__)
#define RSDFT_API __attribute__((visibility("default")))
#else
#define RSDFT_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Precision modes for rsdft_gpu_set_precision(). */
#define RSDFT_PREC_FP64       0  /* FP64 storage, FP64 operators and GEMMs            */
#define RSDFT_PREC_MIXED_FP32 1  /* FP32 storage/operators, FP32 GEMMs, FP64 small LA */
#define RSDFT_PREC_MIXED_BF16 2  /* FP32 storage/operators, BF16x3 split MFMA GEMMs   */

/* Grid, FD kinetic coefficients and the cell volume element.
 * Replaces the set-up done by Init_RgridSol (rgrid_sol_module.f90:63-76),
 * init_kinetic (kinetic_module.f90:20-153) and init_bcset (bc_module.f90);
 * the periodic halo becomes in-device modular indexing. */
RSDFT_API int rsdft_gpu_init(int N1, int N2, int N3, const double *Hgrid, double dV, int Md,
                   const double *coef_lap, double coef_lap0);

/* Local potential Vloc(ML_0:ML_1, s) (localpot_module.f90:14-21). */
RSDFT_API int rsdft_gpu_set_vloc(const double *vloc, int s);

/* KB projector tables as prep_ps_nloc2 leaves them
 * (ps_nloc2_variables.f90:5-27, ps_nloc2_module.f90:912-927). */
RSDFT_API int rsdft_gpu_set_projectors(int nzlma, int maxmjj, const int *MJJ, const int *JJP,
                             const double *uVk, const int *iuV);

/* Position-derivative tables for the nonlocal force: amap(nzlma) 1-based atom
 * of each projector, MJJ_MAP(nzlma), JJ_MAP(MMJJ,nzlma) 1-based grid index of
 * every sphere point, duVdR(3,MMJJ,nzlma) = d uV / d R_atom
 * (ps_nloc2_module.f90:1489-1587). */
RSDFT_API int rsdft_gpu_set_projector_derivatives(int natom, const int *amap, int mmjj,
                                        const int *MJJ_MAP, const int *JJ_MAP,
                                        const double *duVdR);

/* RSDFT_PREC_* for every later call. */
RSDFT_API int rsdft_gpu_set_precision(int mode);

/* Ncg (cg_module.f90:23, keyword NCG) and mloop (cgpc_module.f90:95-96). */
RSDFT_API int rsdft_gpu_set_solver_params(int ncg, int mloop);

/* How precond_hprsdft (cgpc_hprsdft_module.f90:126-263) is evaluated:
 * 1 (default) s-step form -- the same mloop CG steps replayed in the
 * coefficient space of a Chebyshev Krylov basis from one fused Gram pass
 * (mloop <= 4); 0 -- the step-by-step recurrence. */
RSDFT_API int rsdft_gpu_set_precond_mode(int mode);

/* Eigensolver of the subspace_diag step (subspace_diag_la_module.f90:159-169):
 * 1 (default) quadratically convergent refinement from the identity when
 * Hsub is nearly diagonal, otherwise mode 3; 3 -- Householder reduction to
 * tridiagonal form in one cooperative launch with Hsub resident in LDS, the
 * tridiagonal eigenpairs by multisection on Sturm counts + inverse iteration
 * (Gram-Schmidt inside clusters, one Newton-Schulz orthogonalisation step;
 * rocsolver_dstedc above n = 512), then the blocked back-transformation
 * (rocsolver_dormtr above n = 512); 2 -- the same with rocsolver_dstedc always;
 * 0 -- always rocsolver_dsyevd (also when Hsub does not fit the LDS). */
RSDFT_API int rsdft_gpu_set_eigensolver(int mode);

/* How rsdft_gpu_sweep runs gram_schmidt + subspace_diag after CG
 * (sweep_module.f90:147-158): 1 (default) fused -- R^-1 of the Cholesky
 * factor of dV X^T X, Hsub = R^-T (0.5 dV (X^T HX + HX^T X)) R^-1 from the
 * H x that CG leaves (recomputed only for bands with a rejected step) and
 * one rotation psi <- X R^-1 V; equal to the two steps in exact arithmetic,
 * one H psi and one rotation fewer.  0 -- the two steps as in the reference. */
RSDFT_API int rsdft_gpu_set_sweep_mode(int mode);

/* hamiltonian(k,s,tpsi,htpsi,n1,n2,ib1,ib2) (hamiltonian_module.f90:20-72):
 * htpsi = T tpsi + Vloc tpsi + V_NL tpsi (htpsi overwritten).  Gamma only:
 * k must be 1, s must be 1. */
RSDFT_API int rsdft_gpu_hamiltonian(int k, int s, const double *tpsi, double *htpsi,
                          int n1, int n2, int ib1, int ib2);

/* Allocation-free hamiltonian: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same operator as
 * rsdft_gpu_hamiltonian, always in FP64, on DEVICE arrays d_tpsi / d_htpsi
 * (ML x nb, band-major); the projection coefficients live in a static 16 MiB
 * device array per GPU, so nb * nzlma must not exceed 2,097,152 (else
 * kErrArg).  Calls no hipMalloc/hipFree; synchronises the library stream
 * before returning.  Calls that overlap on one GPU share that array and race. */
RSDFT_API int rsdft_gpu_hamiltonian_noalloc(int k, int s, const double *d_tpsi, double *d_htpsi,
                          int n1, int n2, int ib1, int ib2);

/* op_kinetic (kinetic_module.f90:186-202 -> kinetic_sol_module.f90:20-149):
 * OVERWRITES htpsi. */
RSDFT_API int rsdft_gpu_op_kinetic(int k, const double *tpsi, double *htpsi,
                         int n1, int n2, int ib1, int ib2);

/* op_localpot(s,mm,nn,f,vf) (localpot_module.f90:27-81): vf += Vloc*f. */
RSDFT_API int rsdft_gpu_op_localpot(int s, int mm, int nn, const double *f, double *vf);

/* op_nonlocal (nonlocal_module.f90:32-48 -> ps_nloc2_op_module.f90:136-270):
 * htpsi += V_NL tpsi. */
RSDFT_API int rsdft_gpu_op_nonlocal(int k, int s, const double *tpsi, double *htpsi,
                          int n1, int n2, int ib1, int ib2);

/* Device mirror of unk(n1:n2, 1:MB) (wf_module.f90:118).  dir 0: device to
 * host, dir 1: host to device. */
RSDFT_API int rsdft_gpu_wf_attach(double *unk, int ML0, int MB);
RSDFT_API int rsdft_gpu_wf_sync(int dir);
RSDFT_API int rsdft_gpu_wf_detach(void);

/* conjugate_gradient(n1,n2,MB,k,s,unk,esp,res) (cg_module.f90:54-94,
 * conjugate_gradient_1 :97-430 with precond_hprsdft
 * cgpc_hprsdft_module.f90:126-348), all bands batched on the device. */
RSDFT_API int rsdft_gpu_conjugate_gradient(int n1, int n2, int MB, int k, int s,
                                 double *unk, double *esp, double *res);

/* gram_schmidt(n0,n1,k,s) (gram_schmidt_module.f90:28-55 ->
 * gram_schmidt_t_module.f90:84-157): orthonormalises all MB bands of unk
 * with respect to dV (the reference ignores n0, n1 likewise). */
RSDFT_API int rsdft_gpu_gram_schmidt(double *unk, int ML0, int MB, int n0, int n1);

/* subspace_diag(k,s) (subspace_diag_module.f90:17-25 ->
 * subspace_diag_la_module.f90:18-232): Rayleigh-Ritz in place, esp(1:MB)
 * ascending. */
RSDFT_API int rsdft_gpu_subspace_diag(double *unk, double *esp, int ML0, int MB);

/* nsweep iterations of the calc_sweep body (sweep_module.f90:147-158):
 * CG -> GS -> subspace_diag, with esp/res of the last sweep returned. */
RSDFT_API int rsdft_gpu_sweep(double *unk, double *esp, double *res, int ML0, int MB, int nsweep);

/* calc_density, NCPP (density_module.f90:129-136): rho = sum occ psi^2. */
RSDFT_API int rsdft_gpu_calc_density(const double *unk, int ML0, int MB, const double *occ, double *rho);

/* ---- SCF iteration around the sweep (scf_module.f90:146-366), Nspin = 1,
 * Gamma point, real wavefunctions ---------------------------------------- */

/* calc_hartree (hartree_module -> hartree_sol_module.f90:19-122): Vh of the
 * density rho(ML) by FFT Poisson on the G sphere of ggrid_module.f90:55-95,
 * E_h = dV/2 sum Vh rho (may be null). */
RSDFT_API int rsdft_gpu_calc_hartree(const double *rho, double *vh, double *E_h);

/* calc_xc for the Perdew-Zunger 81 LDA (xc_ldapz81_module.f90:17-75, MSP=1):
 * Vxc(ML) and E_xc[3] = (Exc, Ex, Ec) (may be null). */
RSDFT_API int rsdft_gpu_calc_xc(const double *rho, double *vxc, double *E_xc);

/* calc_fermi with Methfessel-Paxton smearing (fermi_mp_module.f90:96-287),
 * one spin, one k point: occ(nb) (2 per filled band) and the Fermi level.
 * Host arithmetic; needs no rsdft_gpu_init. */
RSDFT_API int rsdft_gpu_calc_fermi(const double *esp, int nb, double nelectron, double ekbt,
                          int kinteg, double *occ, double *efermi);

/* init_scf / init_mixing (scf_module.f90:100-145, mixing_module.f90:69-140)
 * for imix = 10 (Pulay on rho): the ionic local potential Vion(ML), the
 * input density rho(ML), the electron count, mmix, beta_mix, ekbt and the
 * Methfessel-Paxton order.  Sets the hamiltonian's Vloc = Vion + Vh + Vxc. */
RSDFT_API int rsdft_gpu_scf_init(const double *vion, const double *rho, double nelectron,
                          int mmix, double beta, double ekbt, int kinteg);

/* One iteration of the scf loop body (scf_module.f90:163-339 with Ndiag = 1,
 * second_diag = 2): subspace_diag, conjugate_gradient, gram_schmidt,
 * esp_calc, calc_fermi, Harris energy, calc_density, calc_hartree, calc_xc,
 * calc_total_energy, calc_sqerr_mixing, perform_mixing, normalize_density
 * and the new Vloc.  unk(ML0,MB) in/out (attached or staged), esp/occ/res(MB)
 * out, info[10] = Etot, Ehwf, Eeig, Eloc, Eion, E_H, E_xc, Efermi,
 * sqerr_rho, sqerr_v. */
RSDFT_API int rsdft_gpu_scf_iterate(double *unk, int ML0, int MB, double *esp, double *occ,
                          double *res, double *info);

/* The SCF fields after the last step: rho, Vloc, Vh, Vxc (ML each; any may be
 * null) -- the module arrays density_module::rho, localpot_module::Vloc,
 * hartree_variables::Vh, xc_module::Vxc. */
RSDFT_API int rsdft_gpu_scf_get_fields(double *rho, double *vloc, double *vh, double *vxc);

/* calc_force_ps_nloc2 for real psi on one rank (ps_nloc2_module.f90:1281-1811):
 * force2(3, natom) = -dE/dR of the nonlocal band energy at fixed psi. */
RSDFT_API int rsdft_gpu_force_ps_nloc2(const double *unk, int ML0, int MB, const double *occ,
                             double *force2);

/* Backward entry of E = sum_n occ_n <psi_n|H|psi_n> dV at fixed psi:
 * *E, dEdV(ML) = dV sum_n occ_n psi_n^2 (the density_module.f90:129-136
 * density times dV) and dEdR(3, natom) = -force2.  Any output may be NULL. */
RSDFT_API int rsdft_gpu_band_energy_grad(const double *unk, int ML0, int MB, const double *occ,
                               double *E, double *dEdV, double *dEdR);

/* test_hpsi2 (test_hpsi2_module.f90:16-143): times nrep applications to the
 * attached wavefunctions on the library's stream; times_ms[8] = mean ms per
 * all-band application of [0] hamiltonian, [1] op_kinetic, [2] op_localpot,
 * [3] op_nonlocal, [4] the stencil kernel as the hamiltonian runs it
 * (kinetic + local accumulated onto the nonlocal term), [5] the projection
 * kernel, [6] the nonlocal term stage (written at every grid point in grid
 * order; scatter-added at the sphere points when a point lies in more than 4
 * projector spheres), [7] a plain copy of all bands (the bandwidth
 * reference). */
RSDFT_API int rsdft_gpu_test_hpsi(int nrep, double *times_ms);

/* Diagnostics of the subspace eigensolve (the dsyevd call of
 * subspace_diag_la_module.f90:159-169): eigenpairs of the symmetric n x n
 * host matrix H (column-major, full storage) by the eigensolver `mode` of
 * rsdft_gpu_set_eigensolver; w[n] ascending, V (n x n, may be NULL) with
 * V(j,j) >= 0, *ms (may be NULL) the device time of the solve. */
RSDFT_API int rsdft_gpu_test_eigh(const double *H, int n, int mode, double *w, double *V,
                                  double *ms);

/* ---- on-disk formats (Fortran sequential unformatted, one process,
 * periodic grid N1 x N2 x N3, LL2 = natural order with 0-based indices).
 * None of these needs rsdft_gpu_init, except that an unk which is the
 * attached array is written from / read into its device mirror. ---- */

/* simple_wf_io_write (io_write_module.f90:19-264), IO_ctrl = 0, real
 * wavefunctions, Gamma, one spin: OC 1, 3 real(8) / 4, 5 real(4) records in
 * the new format (-1 marker, sizes, MK MS MMBZ, IO_ctrl OC TYPE_WF, LL2, occ,
 * aa bb kbb, np(0:7) and the four distributions); OC 11, 13 the old format.
 * Bands MBwr1..MBwr2 (all when out of range, as the reference resets them).
 * aa, bb: 3x3 column-major (may be NULL: zeros). */
RSDFT_API int rsdft_gpu_write_wf(const char *file, int N1, int N2, int N3, const double *unk,
                                 int ML0, int MB, int OC, int MBwr1, int MBwr2,
                                 const double *occ, const double *aa, const double *bb);

/* simple_wf_io_read (io_read_module.f90:15-521), SYStype = 0, IO_ctrl = 0:
 * either format, real(8) or real(4) records, values mapped through the
 * file's LL2.  unk(ML0, MB): bands beyond the file's range are left as they
 * are, file bands beyond MB are skipped; occ (may be NULL) gets
 * min(MB, file MB) occupations; *mb_file (may be NULL) the file's MB.  A
 * different grid, k-points, spins or complex data are refused. */
RSDFT_API int rsdft_gpu_read_wf(const char *file, int N1, int N2, int N3, double *unk, int ML0,
                                int MB, double *occ, int *mb_file);

/* write_data's density/potential file (io_module.f90:96-200), versioned
 * format: version, sizes, LL2, lattice, atoms (aa_atom(3,natom),
 * ki_atom(natom), zn_atom(nelement)), Fermi energy, then per spin rho, Vloc,
 * (spin 1) Vh, Vxc.  rho, vloc, vxc: (ML0, MSP); vh: (ML0). */
RSDFT_API int rsdft_gpu_write_vrho(const char *file, int N1, int N2, int N3, int ML0, int MSP,
                                   const double *rho, const double *vloc, const double *vh,
                                   const double *vxc, const double *aa, int nelement, int natom,
                                   const double *aa_atom, const int *ki_atom, const int *zn_atom,
                                   double efermi);

/* read_data's density/potential part (io_module.f90:286-409): versioned and
 * unversioned files; any output may be NULL (skipped); *efermi = 0 for an
 * unversioned file. */
RSDFT_API int rsdft_gpu_read_vrho(const char *file, int N1, int N2, int N3, int ML0, int MSP,
                                  double *rho, double *vloc, double *vh, double *vxc,
                                  double *efermi);

/* Measurement: device time in ms (mean over nrep) of the tall GEMMs of one
 * sweep on the attached wavefunctions, in the working precision:
 * [0] the overlap dV X^T X (gram_schmidt_t_module.f90:187-199),
 * [1] X^T (H X) as the subspace matrix is formed (one triangle of tiles),
 * [2] the rotation X U (subspace_rotv_sl_module.f90:24-135).  No reference
 * counterpart (the reference times whole routines only). */
RSDFT_API int rsdft_gpu_test_gemm(int nrep, double *times_ms);

/* Per-phase timers in milliseconds since the last reset (watch_module.f90
 * time_hmlt/time_cgpc analogue): [0] hamiltonian, [1] preconditioner,
 * [2] CG total, [3] Gram-Schmidt, [4] subspace_diag, [5] sweep total,
 * [6] number of band-hamiltonian applications (bands a masked launch
 * skips are not counted), [7] number of subspace eigensolves done by
 * refinement, [8] number of bands whose CG step was rejected
 * (cg_module.f90:361-364) in the last CG call, [9] number of band
 * preconditionings done by the step-by-step recurrence because the s-step
 * replay could not resolve them (since the last reset).  n <= 10.  The hamiltonian
 * and preconditioner timers run inside the CG, Gram-Schmidt and sweep
 * timers (they overlap them, as the reference's time_hmlt does). */
RSDFT_API int rsdft_gpu_get_timers(double *t, int n);
RSDFT_API int rsdft_gpu_reset_timers(void);

/* ---- complex wavefunctions at one k point (orthorhombic cell) ----------
 * psi = exp(ik.r) u; the arrays below hold u as complex(8) (interleaved
 * real/imaginary doubles), band-major (ML0, MB) as the reference's
 * complex unk(n1:n2, MB).  FP64 throughout.
 *
 * The k point's kinetic terms and Bloch-phased projector tables: const_k2 =
 * |k|^2/2 and coef_nabk(3, Md) = nab(m)/h_d * k_d (kinetic_module.f90:
 * 118-149: zcoef_kin(+m) = coef_lap - i coef_nabk), uVk complex(maxmjj,
 * nzlma) = uV exp(-i k.x) summed over the images of a point
 * (prep_uvk_ps_nloc2, ps_nloc2_module.f90:1128-1190), MJJ / JJP / iuV as
 * for rsdft_gpu_set_projectors.  The real local potential and the real FD
 * Laplacian are the context's (rsdft_gpu_init, rsdft_gpu_set_vloc). */
RSDFT_API int rsdft_gpu_set_kpoint(double const_k2, const double *coef_nabk, int nzlma,
                                   int maxmjj, const int *MJJ, const int *JJP,
                                   const double *uVk, const int *iuV);
/* hamiltonian(k,s,tpsi,htpsi,n1,n2,ib1,ib2) for complex(8) tpsi/htpsi
 * (hamiltonian_module.f90:20-72, complex branch: op_kinetic_sol with
 * zcoef_kin, op_localpot, op_nonlocal with uVk).  k must be 1. */
RSDFT_API int rsdft_gpu_zhamiltonian(int k, int s, const double *tpsi, double *htpsi, int n1,
                                     int n2, int ib1, int ib2);
/* conjugate_gradient_1, complex branch (cg_module.f90:97-430 with the zhegv
 * 2x2 step of :299-318, precond_hprsdft cgpc_hprsdft_module.f90:126-263):
 * unk complex(ML0, MB) in/out, esp and res out. */
RSDFT_API int rsdft_gpu_zconjugate_gradient(double *unk, double *esp, double *res, int ML0,
                                            int MB);
/* gram_schmidt (gram_schmidt_t_module.f90:84-316) on complex unk. */
RSDFT_API int rsdft_gpu_zgram_schmidt(double *unk, int ML0, int MB);
/* subspace_diag (subspace_diag_la_module.f90:85-206, zheevd) on complex unk. */
RSDFT_API int rsdft_gpu_zsubspace_diag(double *unk, double *esp, int ML0, int MB);
/* nsweep iterations of conjugate_gradient -> gram_schmidt -> subspace_diag
 * (sweep_module.f90:147-158) at the k point. */
RSDFT_API int rsdft_gpu_zsweep(double *unk, double *esp, double *res, int ML0, int MB,
                               int nsweep);

/* Release every device resource. */
RSDFT_API int rsdft_gpu_finalize(void);

/* Message of the last failing call (static storage). */
RSDFT_API const char *rsdft_gpu_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RSDFT_GPU_H */
