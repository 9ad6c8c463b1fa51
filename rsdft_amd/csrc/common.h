// Internal definitions of the MI355X Kohn-Sham library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace rsdft {

constexpr int kMaxMd = 8;
// hamiltonian stencil layouts: FP32 moves chunks of 2 consecutive x points per
// thread, FP64 two x points tpr apart (strided); workgroup size limits
constexpr int kHpsiPx32 = 2, kHpsiPx64 = 2;
constexpr int kHpsiThreads32 = 1024, kHpsiThreads64 = 512;
// H psi: the nonlocal term is written for every grid point of each band
// (k_nl_fill, ranges of kFillRange points, at most kFillMaxM projector groups
// covering a point and kFillMaxG groups touching a range) and the stencil
// accumulates onto it
constexpr int kFillRange = 1024, kFillMaxM = 4, kFillMaxG = 64;
// LDS bank period in bytes.  The stencils' LDS row pitches are congruent to
// the row's own extent modulo this, so the lanes of a wave that run past the
// end of one row continue the bank sequence in the next instead of landing on
// the banks its first lanes use.  256 (64 banks of 4 bytes) measured faster on
// gfx950 than 128: Si216 FP64 preconditioner 7.27 vs 7.64 ms per sweep, FP32
// stencil 0.502 vs 0.512 ms, against 7.72 and 0.523 ms with the unpadded rows
constexpr int kLdsPassBytes = 256;

enum Status {
    kOK = 0,
    kErrNotInit = 1,
    kErrArg = 2,
    kErrHip = 3,
    kErrBlas = 4,
    kErrSolver = 5,
    kErrState = 6,
};

int set_error(int code, const char *fmt, ...);
const char *last_error();

#define RS_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::rsdft::set_error(::rsdft::kErrHip, "%s:%d %s -> %s", __FILE__,    \
                                      __LINE__, #call, hipGetErrorString(e_));         \
    } while (0)

#define RS_BLAS(call)                                                                  \
    do {                                                                               \
        rocblas_status s_ = (call);                                                    \
        if (s_ != rocblas_status_success)                                              \
            return ::rsdft::set_error(::rsdft::kErrBlas, "%s:%d %s -> rocblas %d",     \
                                      __FILE__, __LINE__, #call, (int)s_);             \
    } while (0)

#define RS_TRY(call)                                                                   \
    do {                                                                               \
        int r_ = (call);                                                               \
        if (r_ != 0) return r_;                                                        \
    } while (0)

// Device allocation that frees itself.
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t nbytes);   // grow-only
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Grid + FD description (periodic orthorhombic cell, Gamma point).
struct GridDesc {
    int N1 = 0, N2 = 0, N3 = 0, Md = 0;
    long ML = 0;
    double h[3] = {0, 0, 0};
    double dV = 0.0;
    double coef_lap0 = 0.0;
    double coef_lap[3][kMaxMd] = {};   // [axis][m-1]
};

// Sparse projector operator in two orientations.
//   rows  : projector lma -> its grid points (CSR over lma), used to project
//   cols  : grid point -> (lma, value) (CSR over grid points), used to apply
// Kleinman-Bylander projector tables.  Projectors that share a sphere (the
// s, p, ... channels of one atom: consecutive lma in the reference's order)
// form a group with one union point list and gp value slots per point, so a
// grid value is gathered once for all channels of its atom.  Groups of one
// colour have disjoint point lists, so the apply can add to the output
// without atomics, colour by colour.
constexpr int kGroupMax = 16;
struct NonlocalTables {
    int nz = 0;
    int ng = 0;        // groups
    int gp = 1;        // value slots per point (power of two >= largest group)
    long npts = 0;     // sum of union point counts
    DevMem g_lma;      // [ng*gp] lma of each slot, -1 for an unused slot
    DevMem g_ptr, g_idx;        // CSR over groups of the union points (0-based grid index)
    DevMem g_val64, g_val32;    // [npts*gp] uV of each slot at each point (0 outside its sphere)
    DevMem g_order;             // groups ordered by colour
    std::vector<int> color_ptr; // colour k: g_order[color_ptr[k] .. color_ptr[k+1])
    // grid-order fill of the nonlocal term (H psi): the grid cut into
    // ranges of kFillRange points, the groups touching each range (CSR
    // fill_rptr/fill_rgrp), and per grid point its membership count and the
    // range-local slot of each membership (fill_meta: bits 0-2 count, 6 bits
    // per slot from bit 3) with the uV values of membership m in
    // fill_uv[(m*ML + i)*gp + s].  fill_mm = 0: the layout does not fit
    // (more than kFillMaxM memberships, kFillMaxG groups per range, ...)
    int fill_mm = 0, fill_gmax = 0, nrange = 0;
    DevMem fill_rptr, fill_rgrp, fill_meta, fill_uv, fill_uv64;   // uV in FP32 / FP64
    DevMem scale64;             // iuV(lma) * dV
    // force tables
    int natom = 0;
    long nnzd = 0;
    DevMem d_ptr, d_idx, d_val;   // CSR over lma of MJJ_MAP points, 3 values each (FP64)
    std::vector<int> amap;        // 0-based atom of each lma (host)
    DevMem amap_dev;
};

// Stencil launch geometry chosen for the grid.
struct StencilGeom {
    int px = 4;   // x points per thread
    int ty = 1;   // y rows per workgroup
    int tz = 1;   // z planes per workgroup (z-march length)
    int threads = 64;
    int tiles = 1;  // (N2/ty)*(N3/tz)
    int hm = 1;     // y-halo items per thread (ceil(2*Md/ty))
    size_t lds = 0; // dynamic LDS bytes
};

// CG per-band scalars (FP64 on device).
struct CGScalars {
    DevMem buf;       // all double arrays below, MB each
    DevMem flags;     // int arrays
    double *E, *E1, *rb, *res, *gkgk, *bk, *u1, *u2, *W, *rr0, *rr1, *pa, *pb;
    int *active, *accept, *pc_active;
    int *rejected;    // 1 when a step of this band was rejected in the last CG call
    int *pc_ill;      // s-step replay unresolved: the step-by-step preconditioner runs
    int *pc_ok;       // active and resolved by the s-step replay
    int nb = 0;       // bands of the last CG call
    int ensure(int MB);
};

struct Timers {
    double ms[8] = {};
};

// Timer slots of rsdft_gpu_get_timers (the reference's watch_module groups:
// hamiltonian, preconditioner, CG, Gram-Schmidt, subspace diag, whole sweep)
enum { kTHam = 0, kTPc = 1, kTCg = 2, kTGs = 3, kTSd = 4, kTSweep = 5 };

// HIP event pairs on the library stream, summed into Timers at flush().
// begin() returns a handle that end() closes, so phases may nest.
struct PhaseEvents {
    std::vector<hipEvent_t> a, b;
    std::vector<int> slot;
    int begin(int s, int *handle);
    int end(int handle);
    int flush();   // the stream must be synchronised
};

struct Context {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    GridDesc g;
    StencilGeom geom_h;    // Md-order hamiltonian stencil, FP64 (2 points per thread)
    StencilGeom geom_h32;  // Md-order hamiltonian stencil, FP32 (4 points per thread)
    StencilGeom geom_pc;   // 7-point preconditioner stencil
    int prec = 0;
    int ncg = 3, mloop = 3;
    int pc_mode = 1;   // 0: step-by-step preconditioner recurrence, 1: s-step form
    size_t max_lds = 64 * 1024;   // LDS bytes one workgroup may allocate on this device
    int eig_mode = 1;  // 0: dsyevd, 1: refinement when Hsub is nearly diagonal, else 3; 2: own tridiagonal reduction + dstedc; 3: own reduction + bisection / inverse iteration
    int sweep_mode = 1;  // 1: Rayleigh-Ritz through the Cholesky factor and CG's H x, 0: GS then subspace_diag
    long eig_refined = 0;
    // eig_mode 1: after a refinement attempt fails (the Ritz vectors still
    // move) the next refine_skip eigensolves go to the tridiagonal path
    // directly; the skip doubles per consecutive failure (at most 8) and
    // resets on a success or a change of eigensolver mode
    int refine_skip = 0, refine_fails = 0;

    DevMem vloc64, vloc32;
    NonlocalTables nl;

    // wavefunction mirror and work arrays, element type by precision
    double *host_unk = nullptr;
    int MB = 0;
    int wf_prec = -1;
    DevMem psi, hx, hp, gk, pgk, pk, rk, ppc, qpc, tmp;
    DevMem coefC;      // projection coefficients [lma][band]
    DevMem part;       // reduction partial sums
    DevMem stage;      // host<->device staging (FP64)
    DevMem opbuf;      // working-precision staging of the op_* entry points
    DevMem mat, mat2, mat3, mat4, eval, einfo, evwork;   // nb x nb FP64 dense work
    DevMem matT;       // nb x nb in working precision
    DevMem dots;       // generic per-band result buffer (FP64)
    DevMem pc_coef;    // s-step preconditioner coefficients [band][mloop]
    CGScalars cg;
    Timers timers;
    PhaseEvents *pe = nullptr;   // events of the running entry point, null: untimed
    long hpsi_bands = 0;         // bands of unmasked H psi launches
    DevMem hpsi_count;           // device counter: bands of masked H psi launches
    DevMem pc_fallback;          // device counters: band preconditionings by the recurrence
                                 // where the s-step replay was unresolved (cumulative, last call)
    // BF16 split-GEMM scratch
    DevMem gemm_part;
};

Context &ctx();

// Phase timing from inside the drivers: no-ops when no entry point is timing.
int phase_begin(int slot, int *handle);
int phase_end(int handle);
// Adds the number of set flags of band_active[0..nb) to the masked H psi count.
int count_active_bands(const int *band_active, int nb);

// ---- launchers (implemented in the .hip files) ----------------------------
// geometry
StencilGeom choose_geom(int N1, int N2, int N3, int md, int px, int elem_bytes);
// Tile of the Md-order hamiltonian stencil: px points per thread, LDS row of
// N1 + 2*pad elements, two LDS plane buffers.
StencilGeom choose_hpsi_geom(int N1, int N2, int N3, int md, int px, int pad, int elem_bytes,
                             int max_threads);

// Hamiltonian pieces. flags bit0 kinetic, bit1 local, bit2 nonlocal, bit3 accumulate.
enum { kKin = 1, kLoc = 2, kNl = 4, kAcc = 8 };
// coef: caller-supplied room for nb * nz projection coefficients (null: the
// context's buffer, grown as needed)
template <class T>
int hpsi_apply(const T *in, T *out, int nb, int flags, const int *band_active = nullptr,
               T *coef = nullptr);
template <class T>
int nl_project(const T *in, int nb, T *coef, const int *band_active = nullptr);
// The pieces of one H psi, for timing: the stencil kernel as hpsi_apply runs
// it (with the fill layout: kinetic + local accumulated onto out; else
// kinetic + local), the nonlocal term stage (with the fill layout: the term
// written for every grid point; else scatter-added into out), and a plain
// copy of nb bands.
template <class T>
int hpsi_stencil_only(const T *in, T *out, int nb);
template <class T>
int nl_stage(T *out, int nb, const T *coef);
// kinetic + local stencil alone, overwriting out (nb rows)
template <class T>
int stencil_kinloc(const T *in, T *out, int nb);
// device state of the complex k-point path
void kpoint_release();
template <class T>
int copy_bands(const T *in, T *out, int nb);
// out += V_NL in from the projection coefficients coef[lma][band]
template <class T>
int nl_apply(T *out, int nb, const T *coef, const int *band_active = nullptr);

// preconditioner (precond_hprsdft) for all active bands by the reference's
// step-by-step recurrence: in g, in/out pg.
template <class T>
int precond_batched(const T *g, T *pg, int nb, const int *active);

// s-step form of the same preconditioner (precond_sstep.hip), two passes:
// coefficients d[b][0..mloop) and pgg[b] = (Pg, g) (may be null), with
// ill[b] = 1 (ill may be null) where the coefficient replay cannot resolve
// the recurrence (nearly dependent Krylov vectors); then out = g + sum d_k
// v_k, or with fuse = 1 out = that + beta out (beta == null: overwrite)
template <class T>
int precond_sstep_coeffs(const T *g, int nb, const int *active, double *part, double *dcoef,
                         double *pgg, int *ill);
template <class T>
int precond_sstep_apply(const T *g, T *out, int nb, const int *active, const double *dcoef,
                        int fuse, const double *beta);
size_t precond_sstep_part_elems(int nb);

// reductions: out[k*nb + b] = scale * sum_i a_k[b,i]*b_k[b,i], k < ndot
template <class T>
int band_dots(int ndot, const T *const *a, const T *const *b, int nb, double scale,
              double *out, const int *active = nullptr);

// CG driver (batched conjugate_gradient_1)
template <class T>
int cg_batched(T *x, int nb, double *esp_dev, double *res_dev);

// dense pieces
template <class T>
int gram_schmidt_dev(T *x, int nb);
template <class T>
int subspace_diag_dev(T *x, int nb, double *esp_dev);
// Eigenpairs of the symmetric nb x nb H (destroyed) by the eig_mode solver:
// eigenvalues ascending in D (2 nb doubles of room), eigenvectors with
// V(j,j) >= 0 at *Vec (H or Vbuf); failures flagged in c.einfo[3]
int eigh_sub(double *H, int nb, double *Vbuf, double *D, double **Vec);
// The two steps of one sweep after CG, fused: R^-1 of the Cholesky factor of
// dV x^T x into c.mat, then the Ritz pairs of (x R^-1) from x and hx = H x,
// psi <- x R^-1 V (the Gram-Schmidt and subspace_diag results in exact
// arithmetic, without a second H psi or a second rotation)
template <class T>
int cholesky_inverse_dev(const T *x, int nb);
template <class T>
int rayleigh_ritz_factored_dev(T *x, const T *hx, int nb, double *esp_dev);

// device time (ms) of the tall GEMMs of a sweep on the attached psi:
// [0] x^T x, [1] x^T y (symmetric product), [2] x U
int time_gemms(int nrep, double *ms);

// adjoint
template <class T>
int density_dev(const T *x, int nb, const double *occ_dev, double *rho_dev);
template <class T>
int force_dev(const T *x, int nb, const double *occ_dev, double *force_dev_out);

// SCF iteration (scf.hip): Hartree potential and E_H of a device density,
// LDA V_xc and (E_xc, E_x, E_c), Methfessel-Paxton occupations (host), the
// SCF state (input density, Vion, mixing history) and one iteration on the
// device-resident psi; info = Etot, Ehwf, Eeig, Eloc, Eion, E_H, E_xc,
// Efermi, sqerr_rho, sqerr_v
int hartree_dev(const double *rho, double *vh, double *eh);
int xc_dev(const double *rho, double *vxc, double *e3);
int calc_fermi_host(const double *esp, int nb, double nel, double ekbt, int kinteg, double *occ,
                    double *efermi);
int scf_init(const double *vion, const double *rho, double nel, int mmix, double beta, double ekbt,
             int kinteg);
bool scf_ready();
template <class T>
int scf_iterate_dev(int nb, double *esp_dev, double *res_dev, double *esp, double *occ,
                    double *info);
int scf_fields(double *rho, double *vloc, double *vh, double *vxc);
int scf_release();

// element-wise helpers
template <class T>
int convert_from_f64(const double *src, T *dst, long n);
template <class T>
int convert_to_f64(const T *src, double *dst, long n);

}  // namespace rsdft
