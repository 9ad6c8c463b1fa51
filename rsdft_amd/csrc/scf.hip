// Self-consistent-field iteration around the fixed-potential sweep
// (scf_module.f90:146-366) for one spin channel at the Gamma point with real
// wavefunctions, on the device:
//   * Hartree potential (hartree_sol_module.f90:19-122): the 3-D DFT of the
//     density as three passes of a dense DFT along one axis each (every
//     output a fixed-order FP64 sum over its line, lines staged in LDS with
//     the twiddle table), V_H(G) = 4 pi rho(G) / G^2 on the G sphere of
//     ggrid_module.f90:55-95 (G = 0 and the Nyquist planes dropped), and the
//     inverse passes;
//   * Perdew-Zunger 81 LDA (xc_ldapz81_module.f90:17-210), pointwise;
//   * Methfessel-Paxton occupations (fermi_mp_module.f90:96-307), on the host
//     (a bisection over nb eigenvalues);
//   * Pulay density mixing imix = 10 (mixing_module.f90:69-299,
//     pulay_r2_mixing :452-539) with the beta halving of perform_mixing and
//     normalize_density (density_module.f90:100-111);
//   * total and Harris energies (total_energy_module.f90:56-339).
// Every reduction is a fixed-order two-stage sum, so a run is reproducible.
#include "common.h"
#include "rsdft_gpu.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace rsdft {

// ---------------------------------------------------------------------------
// fixed-order FP64 reductions over grid-sized vectors
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 512, kRedThreads = 256;

enum { kRedDot = 0, kRedSqDiff = 1, kRedDiffDot = 2, kRedSum = 3 };

// part[block] = sum over the block's strided share of
//   kRedDot: a b,  kRedSqDiff: (a - b)^2,  kRedDiffDot: (a - b)(c - d),  kRedSum: a
template <int MODE>
__global__ void __launch_bounds__(kRedThreads)
k_vreduce(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
          const double *__restrict__ d, long n, double *__restrict__ part)
{
    __shared__ double red[kRedThreads];
    double t = 0.0;
    for (long i = blockIdx.x * (long)kRedThreads + threadIdx.x; i < n;
         i += (long)gridDim.x * kRedThreads) {
        if (MODE == kRedDot) t += a[i] * b[i];
        else if (MODE == kRedSqDiff) { const double e = a[i] - b[i]; t += e * e; }
        else if (MODE == kRedDiffDot) t += (a[i] - b[i]) * (c[i] - d[i]);
        else t += a[i];
    }
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kRedThreads)
k_vreduce_final(const double *__restrict__ part, int np, double scale, double *__restrict__ out)
{
    __shared__ double red[kRedThreads];
    double t = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedThreads) t += part[i];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = kRedThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = scale * red[0];
}

// y = sum_k w_k x_k (+ 0), up to kMaxMix vectors; also y = alpha a + beta b
constexpr int kMaxMix = 16;
struct MixCoef {
    const double *x[kMaxMix];
    double w[kMaxMix];
    int n;
};

__global__ void __launch_bounds__(256)
k_vcombine(MixCoef mc, long n, double *__restrict__ y)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double t = 0.0;
        for (int k = 0; k < mc.n; ++k) t += mc.w[k] * mc.x[k][i];
        y[i] = t;
    }
}

// y = a + beta (b - a)   (the mixing step, in the order the reference writes it;
// y may be a)
__global__ void __launch_bounds__(256)
k_vmix(const double *a, const double *b, double beta, long n, double *y)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = a[i] + beta * (b[i] - a[i]);
}

// y = a + b + c
__global__ void __launch_bounds__(256)
k_vsum3(const double *__restrict__ a, const double *__restrict__ b, const double *__restrict__ c,
        long n, double *__restrict__ y)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = a[i] + b[i] + c[i];
}

// y = |y| * s
__global__ void __launch_bounds__(256)
k_vabs_scale(double *__restrict__ y, double s, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = fabs(y[i]) * s;
}

// y *= s
__global__ void __launch_bounds__(256)
k_vscale(double *__restrict__ y, double s, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        y[i] = y[i] * s;
}

static int elem_grid(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

// ---------------------------------------------------------------------------
// Dense DFT along one axis of a complex N3 x N2 x N1 array (x fastest):
//   out[l][k] = sum_j in[l][j] tw[(j k) mod N],  tw[m] = exp(sign 2 pi i m / N)
// A workgroup stages LB lines and the twiddles in LDS.  For the x axis the
// lines are contiguous (consecutive threads walk j); for y and z, LB lines
// with consecutive x (consecutive threads walk the lines), so every global
// access is coalesced.  REAL_IN reads a real array (imaginary part 0),
// REAL_OUT writes only the real part.
// ---------------------------------------------------------------------------
constexpr int kDftThreads = 256;
constexpr int kDftLdsElems = 2048;   // staged complex values per workgroup (32 KB)

struct DftAxis {
    int N;        // line length
    long stride;  // element stride along the line
    int nlines;
    int lb;       // lines per workgroup
    int xfast;    // 1: the x axis (contiguous lines)
    int N1;       // x extent
    long outer;   // stride between blocks of N1 lines (y, z axes)
};

__device__ __forceinline__ long dft_line_base(const DftAxis &A, int l)
{
    if (A.xfast) return (long)l * A.N;
    const int a = l % A.N1, b = l / A.N1;
    return a + (long)b * A.outer;
}

template <int REAL_IN, int REAL_OUT>
__global__ void __launch_bounds__(kDftThreads)
k_dft_axis(const double2 *__restrict__ in, const double *__restrict__ rin, double2 *__restrict__ out,
           double *__restrict__ rout, const double2 *__restrict__ tw, DftAxis A)
{
    extern __shared__ double2 dft_lds[];
    double2 *L = dft_lds;                  // [lb][N]
    double2 *T = dft_lds + A.lb * A.N;     // [N]
    const int l0 = blockIdx.x * A.lb;
    const int nl = min(A.lb, A.nlines - l0);
    const int tot = nl * A.N;
    for (int m = threadIdx.x; m < A.N; m += kDftThreads) T[m] = tw[m];
    for (int t = threadIdx.x; t < tot; t += kDftThreads) {
        int li, j;
        if (A.xfast) { li = t / A.N; j = t - li * A.N; }
        else { j = t / nl; li = t - j * nl; }
        const long g = dft_line_base(A, l0 + li) + (long)j * A.stride;
        L[li * A.N + j] = REAL_IN ? make_double2(rin[g], 0.0) : in[g];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < tot; t += kDftThreads) {
        int li, k;
        if (A.xfast) { li = t / A.N; k = t - li * A.N; }
        else { k = t / nl; li = t - k * nl; }
        const double2 *x = L + li * A.N;
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int j = 0; j < A.N; ++j) {
            const double2 v = x[j], w = T[m];
            re += v.x * w.x - v.y * w.y;
            im += v.x * w.y + v.y * w.x;
            m += k;
            if (m >= A.N) m -= A.N;
        }
        const long g = dft_line_base(A, l0 + li) + (long)k * A.stride;
        if (REAL_OUT) rout[g] = re;
        else out[g] = make_double2(re, im);
    }
}

// V_H(G) = 4 pi / G^2 rho(G) / ML on the G sphere, 0 elsewhere
// (hartree_sol_module.f90:56-73; numpy/FFTW frequency order)
__global__ void __launch_bounds__(256)
k_poisson_g(double2 *__restrict__ c, int N1, int N2, int N3, double b1, double b2, double b3,
            double ecut, double scale)
{
    const long ML = (long)N1 * N2 * N3;
    const int ng1 = (N1 - 1) / 2, ng2 = (N2 - 1) / 2, ng3 = (N3 - 1) / 2;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < ML; i += (long)gridDim.x * 256) {
        const int i1 = (int)(i % N1), i2 = (int)((i / N1) % N2), i3 = (int)(i / ((long)N1 * N2));
        const int f1 = i1 < (N1 + 1) / 2 ? i1 : i1 - N1;
        const int f2 = i2 < (N2 + 1) / 2 ? i2 : i2 - N2;
        const int f3 = i3 < (N3 + 1) / 2 ? i3 : i3 - N3;
        const double g1 = b1 * f1, g2 = b2 * f2, g3 = b3 * f3;
        const double gg = g3 * g3 + g2 * g2 + g1 * g1;
        const bool keep = abs(f1) <= ng1 && abs(f2) <= ng2 && abs(f3) <= ng3 &&
                          gg < ecut - 1e-14 && gg != 0.0;
        if (keep) {
            const double s = scale * (4.0 * M_PI) / gg;
            c[i] = make_double2(c[i].x * s, c[i].y * s);
        } else {
            c[i] = make_double2(0.0, 0.0);
        }
    }
}

// ---------------------------------------------------------------------------
// Perdew-Zunger 81 LDA, unpolarised (calc_ldapz81_x / calc_ldapz81_c with
// rho_up = rho_down = rho / 2): vxc and the per-point energy densities
//   ex[i] = rho e_x,  ec[i] = rho e_c
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_lda_pz81(const double *__restrict__ rho, long n, double *__restrict__ vxc,
           double *__restrict__ ex, double *__restrict__ ec)
{
    constexpr double gam = -0.1423, bet1 = 1.0529, bet2 = 0.3334;
    constexpr double A = 0.0311, B = -0.048, C = 0.002, D = -0.0116;
    constexpr double onethr = 1.0 / 3.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double r = rho[i];
        // exchange: e_x = -3/4 (3 rho / pi)^(1/3), v_x = 4/3 e_x
        const double exd = -0.75 * cbrt(3.0 / M_PI * r);
        double v = (4.0 / 3.0) * exd;
        const double trho = fabs(0.5 * r) + fabs(0.5 * r);
        double e1 = 0.0;
        if (trho > 0.0) {
            const double rs = cbrt(3.0 / (4.0 * M_PI) / trho);
            if (rs >= 1.0) {
                const double sq = sqrt(rs);
                e1 = gam / (1.0 + bet1 * sq + bet2 * rs);
                v += e1 + e1 * onethr * (0.5 * bet1 / sq + bet2) / (1.0 / rs + bet1 / sq + bet2);
            } else {
                const double ln = log(rs);
                e1 = A * ln + B + C * rs * ln + D * rs;
                v += e1 - onethr * (A + C * rs * (1.0 + ln) + rs * D);
            }
        }
        vxc[i] = v;
        ex[i] = r * exd;
        ec[i] = trho * e1;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct ScfState {
    bool ready = false;
    int mmix = 4, count = 0, kinteg = 5;
    double beta = 1.0, ekbt = 1e-5, nel = 0.0, dif0 = 0.0;
    double eh = 0.0, exc = 0.0;
    DevMem rho, vion, vh, vxc, vloc;          // FP64 grid fields
    DevMem xin, xou, xin_b, xou_b;            // mmix x ML mixing history and its backup
    DevMem xold_rho, xold_v;                  // the previous (rho, v) pair
    DevMem rho_out, vh_out, vxc_out, vout, hmix, work;   // per-iteration fields
    DevMem cplx, cplx2, tw_f, tw_b;           // DFT work and twiddles (per axis, forward/back)
    DevMem part, scal;                        // reduction partials and results
    std::vector<int> tw_off;                  // twiddle offsets of the three axes
};

static ScfState *const g_scf = new ScfState();

static int release_scf()
{
    ScfState &S = *g_scf;
    DevMem *all[] = {&S.rho, &S.vion, &S.vh, &S.vxc, &S.vloc, &S.xin, &S.xou, &S.xin_b, &S.xou_b,
                     &S.xold_rho, &S.xold_v, &S.rho_out, &S.vh_out, &S.vxc_out, &S.vout, &S.hmix,
                     &S.work, &S.cplx, &S.cplx2, &S.tw_f, &S.tw_b, &S.part, &S.scal};
    for (DevMem *m : all) m->release();
    S.ready = false;
    return kOK;
}

int scf_release() { return release_scf(); }

// fixed-order reduction of n values to one FP64 scalar on the host
template <int MODE>
static int vreduce(const double *a, const double *b, const double *c, const double *d, long n,
                   double scale, double *out)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    RS_TRY(S.part.ensure(sizeof(double) * kRedBlocks));
    RS_TRY(S.scal.ensure(sizeof(double) * 8));
    hipLaunchKernelGGL((k_vreduce<MODE>), dim3(kRedBlocks), dim3(kRedThreads), 0, C.stream, a, b, c,
                       d, n, S.part.as<double>());
    hipLaunchKernelGGL(k_vreduce_final, dim3(1), dim3(kRedThreads), 0, C.stream, S.part.as<double>(),
                       kRedBlocks, scale, S.scal.as<double>());
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(out, S.scal.p, sizeof(double), hipMemcpyDeviceToHost, C.stream));
    RS_HIP(hipStreamSynchronize(C.stream));
    return kOK;
}

static int vdot(const double *a, const double *b, long n, double scale, double *out)
{
    return vreduce<kRedDot>(a, b, nullptr, nullptr, n, scale, out);
}

static int twiddles_ready()
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const int Ns[3] = {C.g.N1, C.g.N2, C.g.N3};
    const size_t tot = (size_t)Ns[0] + Ns[1] + Ns[2];
    if (S.tw_f.p && S.tw_f.bytes >= tot * sizeof(double2) && S.tw_off.size() == 4 &&
        S.tw_off[3] == (int)tot)
        return kOK;
    std::vector<double2> f(tot), b(tot);
    S.tw_off.assign(4, 0);
    size_t o = 0;
    for (int d = 0; d < 3; ++d) {
        S.tw_off[d] = (int)o;
        const int N = Ns[d];
        for (int m = 0; m < N; ++m) {
            const double t = 2.0 * M_PI * (double)m / (double)N;
            const double cs = std::cos(t), sn = std::sin(t);
            f[o + m] = make_double2(cs, -sn);   // forward: exp(-2 pi i m / N)
            b[o + m] = make_double2(cs, sn);
        }
        o += N;
    }
    S.tw_off[3] = (int)o;
    RS_TRY(S.tw_f.ensure(tot * sizeof(double2)));
    RS_TRY(S.tw_b.ensure(tot * sizeof(double2)));
    RS_HIP(hipMemcpy(S.tw_f.p, f.data(), tot * sizeof(double2), hipMemcpyHostToDevice));
    RS_HIP(hipMemcpy(S.tw_b.p, b.data(), tot * sizeof(double2), hipMemcpyHostToDevice));
    return kOK;
}

// one DFT pass along axis d (0 x, 1 y, 2 z)
template <int RI, int RO>
static int dft_pass(int d, bool forward, const double2 *in, const double *rin, double2 *out,
                    double *rout)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const int N1 = C.g.N1, N2 = C.g.N2, N3 = C.g.N3;
    DftAxis A{};
    A.N1 = N1;
    if (d == 0) {
        A.N = N1; A.stride = 1; A.nlines = N2 * N3; A.xfast = 1; A.outer = 0;
    } else if (d == 1) {
        A.N = N2; A.stride = N1; A.nlines = N1 * N3; A.xfast = 0; A.outer = (long)N1 * N2;
    } else {
        A.N = N3; A.stride = (long)N1 * N2; A.nlines = N1 * N2; A.xfast = 0; A.outer = N1;
    }
    if (A.N + A.N > kDftLdsElems) return set_error(kErrArg, "DFT line of %d points too long", A.N);
    A.lb = std::max(1, std::min(64, (kDftLdsElems - A.N) / A.N));
    const double2 *tw = (forward ? S.tw_f.as<double2>() : S.tw_b.as<double2>()) + S.tw_off[d];
    const size_t lds = sizeof(double2) * ((size_t)A.lb * A.N + A.N);
    const int nblk = (A.nlines + A.lb - 1) / A.lb;
    hipLaunchKernelGGL((k_dft_axis<RI, RO>), dim3(nblk), dim3(kDftThreads), lds, C.stream, in, rin,
                       out, rout, tw, A);
    RS_HIP(hipGetLastError());
    return kOK;
}

// V_H of a device density and E_H = dV/2 sum V_H rho (hartree_sol_module.f90:19-122)
int hartree_dev(const double *rho, double *vh, double *eh)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const GridDesc &g = C.g;
    const long ML = g.ML;
    RS_TRY(twiddles_ready());
    RS_TRY(S.cplx.ensure(sizeof(double2) * ML));
    RS_TRY(S.cplx2.ensure(sizeof(double2) * ML));
    double2 *a = S.cplx.as<double2>(), *b = S.cplx2.as<double2>();
    RS_TRY((dft_pass<1, 0>(0, true, nullptr, rho, a, nullptr)));
    RS_TRY((dft_pass<0, 0>(1, true, a, nullptr, b, nullptr)));
    RS_TRY((dft_pass<0, 0>(2, true, b, nullptr, a, nullptr)));
    const double L[3] = {g.N1 * g.h[0], g.N2 * g.h[1], g.N3 * g.h[2]};
    const int Ns[3] = {g.N1, g.N2, g.N3};
    double bb[3], gcut = 1e300;
    for (int d = 0; d < 3; ++d) {
        bb[d] = 2.0 * M_PI / L[d];
        gcut = std::min(gcut, 0.5 * bb[d] * (2 * ((Ns[d] - 1) / 2) + 1));
    }
    hipLaunchKernelGGL(k_poisson_g, dim3(elem_grid(ML)), dim3(256), 0, C.stream, a, g.N1, g.N2,
                       g.N3, bb[0], bb[1], bb[2], gcut * gcut, 1.0 / (double)ML);
    RS_HIP(hipGetLastError());
    RS_TRY((dft_pass<0, 0>(2, false, a, nullptr, b, nullptr)));
    RS_TRY((dft_pass<0, 0>(1, false, b, nullptr, a, nullptr)));
    RS_TRY((dft_pass<0, 1>(0, false, a, nullptr, nullptr, vh)));
    if (eh) RS_TRY(vdot(vh, rho, ML, 0.5 * g.dV, eh));
    return kOK;
}

// V_xc and (E_xc, E_x, E_c) of a device density (calc_LDAPZ81 with MSP = 1)
int xc_dev(const double *rho, double *vxc, double *e3)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const long ML = C.g.ML;
    RS_TRY(S.work.ensure(sizeof(double) * 2 * ML));
    double *ex = S.work.as<double>(), *ec = ex + ML;
    hipLaunchKernelGGL(k_lda_pz81, dim3(elem_grid(ML)), dim3(256), 0, C.stream, rho, ML, vxc, ex, ec);
    RS_HIP(hipGetLastError());
    double x = 0.0, c = 0.0;
    RS_TRY(vreduce<kRedSum>(ex, nullptr, nullptr, nullptr, ML, C.g.dV, &x));
    RS_TRY(vreduce<kRedSum>(ec, nullptr, nullptr, nullptr, ML, C.g.dV, &c));
    e3[0] = x + c;
    e3[1] = x;
    e3[2] = c;
    return kOK;
}

// ---- Methfessel-Paxton occupations (host) ---------------------------------
// The step function of order kinteg (ff0 of fermi_mp_module.f90:289-307):
//   S(x) = erfc(x)/2 + sum_{n=1..kinteg} A_n H_{2n-1}(x) exp(-x^2),
//   A_n = (-1)^n / (n! 4^n sqrt(pi)), H the physicists' Hermite polynomials.
static double mp_step(double x, int kinteg)
{
    double s = 0.5 * (1.0 - std::erf(x));
    if (kinteg <= 0) return s;
    double an = -1.0 / (4.0 * std::sqrt(M_PI));
    double h0 = 1.0, h1 = 2.0 * x;       // H_{2n-2}, H_{2n-1} for n = 1
    double sum = an * h1;
    for (int n = 2; n <= kinteg; ++n) {
        an = -an / (4.0 * n);
        const double h2 = 2.0 * x * h1 - 2.0 * (2 * n - 3) * h0;   // H_{2n-2}
        const double h3 = 2.0 * x * h2 - 2.0 * (2 * n - 2) * h1;   // H_{2n-1}
        sum += an * h3;
        h0 = h2;
        h1 = h3;
    }
    return s + sum * std::exp(-x * x);
}

// calc_fermi for one spin and one k point of weight 1 (fermi_mp_module.f90:
// 96-287): bisection of the Fermi level on the electron count between the
// lowest and the highest eigenvalue (+ min(100 ekbt, 0.1)), at most 1000
// halvings, ending when the midpoint stops moving; occ = 2 S((e - ef)/ekbt).
int calc_fermi_host(const double *esp, int nb, double nel, double ekbt, int kinteg, double *occ,
                    double *efermi)
{
    if (nb <= 0) return set_error(kErrArg, "calc_fermi: nb=%d", nb);
    if (!(ekbt > 0.0)) return set_error(kErrArg, "calc_fermi: ekbt=%g", ekbt);
    double lo = esp[0], hi = esp[0];
    for (int n = 1; n < nb; ++n) {
        lo = std::min(lo, esp[n]);
        hi = std::max(hi, esp[n]);
    }
    if (lo == hi) {
        lo -= 0.01;
        hi += 0.01;
    }
    hi += std::min(ekbt * 1e2, 0.1);
    std::vector<double> ff(nb, 0.0);
    double prev = 1e10, ef = 0.5 * (lo + hi), count = 0.0;
    bool done = false;
    for (int it = 0; it < 1000; ++it) {
        ef = 0.5 * (lo + hi);
        if (ef == prev) {
            done = true;
            break;
        }
        count = 0.0;
        for (int n = 0; n < nb; ++n) {
            ff[n] = mp_step((esp[n] - ef) / ekbt, kinteg);
            count += 2.0 * ff[n];
        }
        if (count > nel) hi = ef;
        else if (count < nel) lo = ef;
        else {
            done = true;
            break;
        }
        prev = ef;
    }
    if (!done && std::fabs(count - nel) > 1e-10)
        return set_error(kErrSolver, "calc_fermi: the Fermi level did not converge");
    for (int n = 0; n < nb; ++n) occ[n] = 2.0 * ff[n];
    *efermi = ef;
    return kOK;
}

// ---- Pulay mixing of the density (imix = 10) ------------------------------
// Solves the m x m system A b = 1 by LU with partial pivoting (zgesv's
// algorithm; A is real symmetric here).
static int solve_small(std::vector<double> A, int m, std::vector<double> &b)
{
    b.assign(m, 1.0);
    for (int k = 0; k < m; ++k) {
        int p = k;
        for (int i = k + 1; i < m; ++i)
            if (std::fabs(A[i * m + k]) > std::fabs(A[p * m + k])) p = i;
        if (A[p * m + k] == 0.0) return set_error(kErrSolver, "mixing: singular Pulay matrix");
        if (p != k) {
            for (int j = 0; j < m; ++j) std::swap(A[k * m + j], A[p * m + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < m; ++i) {
            const double l = A[i * m + k] / A[k * m + k];
            for (int j = k; j < m; ++j) A[i * m + j] -= l * A[k * m + j];
            b[i] -= l * b[k];
        }
    }
    for (int k = m - 1; k >= 0; --k) {
        double t = b[k];
        for (int j = k + 1; j < m; ++j) t -= A[k * m + j] * b[j];
        b[k] = t / A[k * m + k];
    }
    return kOK;
}

static double *slot(DevMem &h, int k) { return h.as<double>() + (size_t)k * ctx().g.ML; }

// shift the history one slot towards 0 (slot m-1 keeps its values)
static int shift_history()
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const size_t nbyte = sizeof(double) * C.g.ML;
    for (int k = 0; k + 1 < S.mmix; ++k) {
        RS_HIP(hipMemcpyAsync(slot(S.xin, k), slot(S.xin, k + 1), nbyte, hipMemcpyDeviceToDevice,
                              C.stream));
        RS_HIP(hipMemcpyAsync(slot(S.xou, k), slot(S.xou, k + 1), nbyte, hipMemcpyDeviceToDevice,
                              C.stream));
    }
    return kOK;
}

// pulay_r2_mixing (mixing_module.f90:452-539): the new input density in h
static int pulay_r2(const double *f, double beta, double *h)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const long ML = C.g.ML;
    const int m = S.mmix;
    RS_HIP(hipMemcpyAsync(slot(S.xou, m - 1), f, sizeof(double) * ML, hipMemcpyDeviceToDevice,
                          C.stream));
    S.count += 1;
    const int m0 = std::min(S.count, m);
    if (m == 1 || m0 < m) {
        RS_TRY(shift_history());
        hipLaunchKernelGGL(k_vmix, dim3(elem_grid(ML)), dim3(256), 0, C.stream, slot(S.xin, m - 1),
                           slot(S.xou, m - 1), beta, ML, slot(S.xin, m - 1));
        RS_HIP(hipGetLastError());
    } else {
        // A_ij = (R_i, R_j), R = Xou - Xin
        std::vector<double> A((size_t)m * m), b;
        for (int i = 0; i < m; ++i)
            for (int j = i; j < m; ++j) {
                double t = 0.0;
                RS_TRY(vreduce<kRedDiffDot>(slot(S.xou, i), slot(S.xin, i), slot(S.xou, j),
                                            slot(S.xin, j), ML, 1.0, &t));
                A[i * m + j] = A[j * m + i] = t;
            }
        RS_TRY(solve_small(A, m, b));
        double sb = 0.0;
        for (int i = 0; i < m; ++i) sb += b[i];
        for (int i = 0; i < m; ++i) b[i] *= 1.0 / sb;
        RS_TRY(S.work.ensure(sizeof(double) * 2 * ML));
        double *X = S.work.as<double>(), *Y = X + ML;
        MixCoef mx{}, my{};
        mx.n = my.n = m;
        for (int i = 0; i < m; ++i) {
            mx.x[i] = slot(S.xin, i);
            my.x[i] = slot(S.xou, i);
            mx.w[i] = my.w[i] = b[i];
        }
        hipLaunchKernelGGL(k_vcombine, dim3(elem_grid(ML)), dim3(256), 0, C.stream, mx, ML, X);
        hipLaunchKernelGGL(k_vcombine, dim3(elem_grid(ML)), dim3(256), 0, C.stream, my, ML, Y);
        RS_HIP(hipGetLastError());
        RS_TRY(shift_history());
        hipLaunchKernelGGL(k_vmix, dim3(elem_grid(ML)), dim3(256), 0, C.stream, X, Y, beta, ML,
                           slot(S.xin, m - 1));
        RS_HIP(hipGetLastError());
    }
    RS_HIP(hipMemcpyAsync(h, slot(S.xin, m - 1), sizeof(double) * ML, hipMemcpyDeviceToDevice,
                          C.stream));
    return kOK;
}

// perform_mixing for imix = 10 (mixing_module.f90:142-286): h = |mixed density|;
// with beta >= 1 up to ten retries with beta halved while the change ratio
// (h - h_old)^2 / previous exceeds 0.5, the last retry with the best beta
static int perform_mixing(const double *f, const double *vout, double *h)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const long ML = C.g.ML;
    const size_t hist = sizeof(double) * ML * S.mmix;
    RS_HIP(hipMemcpyAsync(S.xin_b.p, S.xin.p, hist, hipMemcpyDeviceToDevice, C.stream));
    RS_HIP(hipMemcpyAsync(S.xou_b.p, S.xou.p, hist, hipMemcpyDeviceToDevice, C.stream));
    const int count_b = S.count;
    double beta = S.beta, beta_min = beta, dif_min = 1e10, dif1 = 0.0;
    const int max_loop = S.beta >= 1.0 ? 11 : 1;
    for (int loop = 1; loop <= max_loop; ++loop) {
        RS_TRY(pulay_r2(f, beta, h));
        RS_TRY(vreduce<kRedSqDiff>(h, S.xold_rho.as<double>(), nullptr, nullptr, ML, 1.0 / ML, &dif1));
        const double diff = S.dif0 != 0.0 ? dif1 / S.dif0 : 0.0;
        if (dif1 < dif_min) {
            beta_min = beta;
            dif_min = dif1;
        }
        if (loop < max_loop) {
            if (diff > 0.5) {
                RS_HIP(hipMemcpyAsync(S.xin.p, S.xin_b.p, hist, hipMemcpyDeviceToDevice, C.stream));
                RS_HIP(hipMemcpyAsync(S.xou.p, S.xou_b.p, hist, hipMemcpyDeviceToDevice, C.stream));
                S.count = count_b;
                beta *= 0.5;
            } else {
                break;
            }
            if (loop == max_loop - 1) beta = beta_min;
        }
    }
    S.dif0 = dif1;
    hipLaunchKernelGGL(k_vabs_scale, dim3(elem_grid(ML)), dim3(256), 0, C.stream, h, 1.0, ML);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(S.xold_rho.p, h, sizeof(double) * ML, hipMemcpyDeviceToDevice, C.stream));
    RS_HIP(hipMemcpyAsync(S.xold_v.p, vout, sizeof(double) * ML, hipMemcpyDeviceToDevice, C.stream));
    return kOK;
}

// the local potential of the hamiltonian from the SCF fields: FP64 and the
// FP32 copy the mixed-precision operators read
static int install_vloc()
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const long ML = C.g.ML;
    hipLaunchKernelGGL(k_vsum3, dim3(elem_grid(ML)), dim3(256), 0, C.stream, S.vion.as<double>(),
                       S.vh.as<double>(), S.vxc.as<double>(), ML, S.vloc.as<double>());
    RS_HIP(hipGetLastError());
    RS_TRY(C.vloc64.ensure(sizeof(double) * ML));
    RS_TRY(C.vloc32.ensure(sizeof(float) * ML));
    RS_HIP(hipMemcpyAsync(C.vloc64.p, S.vloc.p, sizeof(double) * ML, hipMemcpyDeviceToDevice,
                          C.stream));
    RS_TRY(convert_from_f64<float>(C.vloc64.as<double>(), C.vloc32.as<float>(), ML));
    return kOK;
}

// potentials of the current input density: V_H, V_xc, Vloc
static int potentials_of_rho()
{
    ScfState &S = *g_scf;
    double e3[3];
    RS_TRY(hartree_dev(S.rho.as<double>(), S.vh.as<double>(), &S.eh));
    RS_TRY(xc_dev(S.rho.as<double>(), S.vxc.as<double>(), e3));
    S.exc = e3[0];
    return install_vloc();
}

int scf_init(const double *vion, const double *rho, double nel, int mmix, double beta, double ekbt,
             int kinteg)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    if (mmix < 1 || mmix > kMaxMix) return set_error(kErrArg, "mmix=%d (1..%d)", mmix, kMaxMix);
    if (!vion || !rho) return set_error(kErrArg, "null Vion or rho");
    const long ML = C.g.ML;
    const size_t f = sizeof(double) * ML;
    DevMem *fields[] = {&S.rho, &S.vion, &S.vh, &S.vxc, &S.vloc, &S.xold_rho, &S.xold_v,
                        &S.rho_out, &S.vh_out, &S.vxc_out, &S.vout, &S.hmix};
    for (DevMem *m : fields) RS_TRY(m->ensure(f));
    for (DevMem *m : {&S.xin, &S.xou, &S.xin_b, &S.xou_b}) RS_TRY(m->ensure(f * mmix));
    S.mmix = mmix;
    S.beta = beta;
    S.ekbt = ekbt;
    S.kinteg = kinteg;
    S.nel = nel;
    S.count = 0;
    S.dif0 = 0.0;
    RS_HIP(hipMemcpyAsync(S.vion.p, vion, f, hipMemcpyHostToDevice, C.stream));
    RS_HIP(hipMemcpyAsync(S.rho.p, rho, f, hipMemcpyHostToDevice, C.stream));
    RS_TRY(potentials_of_rho());
    // init_mixing: the history empty but for the input density in the last slot
    RS_HIP(hipMemsetAsync(S.xin.p, 0, f * mmix, C.stream));
    RS_HIP(hipMemsetAsync(S.xou.p, 0, f * mmix, C.stream));
    RS_HIP(hipMemcpyAsync(slot(S.xin, mmix - 1), S.rho.p, f, hipMemcpyDeviceToDevice, C.stream));
    RS_HIP(hipMemcpyAsync(S.xold_rho.p, S.rho.p, f, hipMemcpyDeviceToDevice, C.stream));
    RS_HIP(hipMemcpyAsync(S.xold_v.p, S.vloc.p, f, hipMemcpyDeviceToDevice, C.stream));
    RS_HIP(hipStreamSynchronize(C.stream));
    S.ready = true;
    return kOK;
}

bool scf_ready() { return g_scf->ready; }

// One iteration of the scf loop body on the device-resident psi (c.psi, nb
// bands in working precision T; the rotations swap c.psi with a work array,
// so psi is looked up after each step): subspace_diag, conjugate_gradient, gram_schmidt,
// esp_calc, calc_fermi, the Harris energy from the input density,
// calc_density, calc_hartree, calc_xc, calc_total_energy, calc_sqerr_mixing,
// perform_mixing, normalize_density and the new Vloc.
// info: Etot, Ehwf, Eeig, Eloc, Eion, E_H, E_xc, Efermi, sqerr_rho, sqerr_v
template <class T>
int scf_iterate_dev(int nb, double *esp_dev, double *res_dev, double *esp, double *occ,
                    double *info)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const long ML = C.g.ML;
    const double dV = C.g.dV;
    RS_TRY(subspace_diag_dev<T>(C.psi.as<T>(), nb, esp_dev));
    RS_TRY(cg_batched<T>(C.psi.as<T>(), nb, esp_dev, res_dev));
    RS_TRY(gram_schmidt_dev<T>(C.psi.as<T>(), nb));
    T *x = C.psi.as<T>();
    // esp_calc (esp_calc_module.f90:36): esp = dV <x|H|x>
    T *hx = C.hx.as<T>();
    RS_TRY(hpsi_apply<T>(x, hx, nb, kKin | kLoc | kNl));
    const T *pa[1] = {x}, *pb[1] = {hx};
    RS_TRY(band_dots<T>(1, pa, pb, nb, dV, esp_dev, nullptr));
    RS_HIP(hipMemcpyAsync(esp, esp_dev, sizeof(double) * nb, hipMemcpyDeviceToHost, C.stream));
    RS_HIP(hipStreamSynchronize(C.stream));
    double ef = 0.0;
    RS_TRY(calc_fermi_host(esp, nb, S.nel, S.ekbt, S.kinteg, occ, &ef));
    double eeig = 0.0;
    for (int n = 0; n < nb; ++n) eeig += occ[n] * esp[n];
    // Harris energy from the input density (calc_with_rhoIN_total_energy)
    const double *rho = S.rho.as<double>(), *vloc = S.vloc.as<double>(), *vion = S.vion.as<double>();
    double eloc_in = 0.0, eion_in = 0.0;
    RS_TRY(vdot(vloc, rho, ML, dV, &eloc_in));
    RS_TRY(vdot(vion, rho, ML, dV, &eion_in));
    const double ehwf = eeig - eloc_in + S.eh + S.exc + eion_in;
    // output density and its energies
    RS_TRY(C.dots.ensure(sizeof(double) * 4 * (size_t)nb));
    double *occ_dev = res_dev + nb;   // the third band-sized slot of c.dots
    RS_HIP(hipMemcpyAsync(occ_dev, occ, sizeof(double) * nb, hipMemcpyHostToDevice, C.stream));
    double *rho_out = S.rho_out.as<double>();
    RS_TRY(density_dev<T>(x, nb, occ_dev, rho_out));
    double eh_out = 0.0, e3[3];
    RS_TRY(hartree_dev(rho_out, S.vh_out.as<double>(), &eh_out));
    RS_TRY(xc_dev(rho_out, S.vxc_out.as<double>(), e3));
    double eloc = 0.0, eion = 0.0;
    RS_TRY(vdot(rho_out, vloc, ML, dV, &eloc));
    RS_TRY(vdot(rho_out, vion, ML, dV, &eion));
    const double etot = eeig - eloc + eh_out + e3[0] + eion;
    hipLaunchKernelGGL(k_vsum3, dim3(elem_grid(ML)), dim3(256), 0, C.stream, vion,
                       S.vh_out.as<double>(), S.vxc_out.as<double>(), ML, S.vout.as<double>());
    RS_HIP(hipGetLastError());
    double sq_rho = 0.0, sq_v = 0.0;
    RS_TRY(vreduce<kRedSqDiff>(rho_out, S.xold_rho.as<double>(), nullptr, nullptr, ML, 1.0 / ML,
                               &sq_rho));
    RS_TRY(vreduce<kRedSqDiff>(S.vout.as<double>(), S.xold_v.as<double>(), nullptr, nullptr, ML,
                               1.0 / ML, &sq_v));
    // mixing, normalize_density, then the potential of the mixed density
    RS_TRY(perform_mixing(rho_out, S.vout.as<double>(), S.hmix.as<double>()));
    double tot = 0.0;
    RS_TRY(vreduce<kRedSum>(S.hmix.as<double>(), nullptr, nullptr, nullptr, ML, dV, &tot));
    RS_HIP(hipMemcpyAsync(S.rho.p, S.hmix.p, sizeof(double) * ML, hipMemcpyDeviceToDevice, C.stream));
    hipLaunchKernelGGL(k_vscale, dim3(elem_grid(ML)), dim3(256), 0, C.stream, S.rho.as<double>(),
                       S.nel / tot, ML);
    RS_HIP(hipGetLastError());
    RS_TRY(potentials_of_rho());
    RS_HIP(hipStreamSynchronize(C.stream));
    const double out[10] = {etot, ehwf, eeig, eloc, eion, eh_out, e3[0], ef, sq_rho, sq_v};
    for (int k = 0; k < 10; ++k) info[k] = out[k];
    return kOK;
}

int scf_fields(double *rho, double *vloc, double *vh, double *vxc)
{
    Context &C = ctx();
    ScfState &S = *g_scf;
    const size_t f = sizeof(double) * C.g.ML;
    if (rho) RS_HIP(hipMemcpyAsync(rho, S.rho.p, f, hipMemcpyDeviceToHost, C.stream));
    if (vloc) RS_HIP(hipMemcpyAsync(vloc, S.vloc.p, f, hipMemcpyDeviceToHost, C.stream));
    if (vh) RS_HIP(hipMemcpyAsync(vh, S.vh.p, f, hipMemcpyDeviceToHost, C.stream));
    if (vxc) RS_HIP(hipMemcpyAsync(vxc, S.vxc.p, f, hipMemcpyDeviceToHost, C.stream));
    RS_HIP(hipStreamSynchronize(C.stream));
    return kOK;
}

template int scf_iterate_dev<double>(int, double *, double *, double *, double *, double *);
template int scf_iterate_dev<float>(int, double *, double *, double *, double *, double *);

}  // namespace rsdft
