// Complex wavefunctions at one k point of an orthorhombic cell: the periodic
// part u of psi = exp(ik.r) u, the hamiltonian -1/2 (grad + ik)^2 + V +
// V_NL(k) on it, and the eigensolver sweep (CG, Gram-Schmidt, Rayleigh-Ritz)
// in complex arithmetic, FP64.
//
// Reference: the complex (non _DRSDFT_) branches of op_kinetic_sol with
// zcoef_kin (kinetic_sol_module.f90:72-125; coefficients kinetic_module.f90:
// 118-149), op_ps_nloc2 with the Bloch-phased uVk (ps_nloc2_op_module.f90:
// 161-175, 255-263, 304-313; tables prep_uvk_ps_nloc2, ps_nloc2_module.f90:
// 1128-1190), precond_hprsdft (cgpc_hprsdft_module.f90:126-263),
// conjugate_gradient_1 with zhegv (cg_module.f90:97-430, :299-318),
// gram_schmidt_t and subspace_diag_la in complex(8).
//
// Layout on the device: complex band b is two rows of ML doubles, its real
// and its imaginary part (rows 2b, 2b+1).  The kinetic Laplacian and Vloc are
// real operators, so the real stencil kernel runs on all 2 nb rows unchanged;
// the k-dependent first-derivative coupling, the Bloch-phased nonlocal term
// and the inner products are kernels of this file; the dense steps are real
// GEMMs on the (2 nb) x ML row matrix with the 2x2 real blocks of the complex
// coefficient matrices (the same flop count as the complex GEMMs), the small
// Hermitian factorisations rocSOLVER's zpotrf / ztrtri / zheevd.
#include "common.h"
#include "rsdft_gpu.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <utility>
#include <vector>

namespace rsdft {

using zc = std::complex<double>;

struct ZNab {
    double c[3][kMaxMd];        // coef_nabk(d, m): nab(m) / h_d * k_d
};

struct KpointState {
    bool set = false;
    double ck2 = 0.0;           // |k|^2 / 2
    ZNab nab{};
    int nz = 0;
    DevMem lptr, lidx, lval;    // per projector: its points (CSR), uVk as (re, im)
    DevMem scale;               // iuV * dV per projector
    int ncov = 0;
    DevMem cptr, cpt, clma, cval;   // per covered point: (projector, uVk) entries
    int MB = 0;
    int nrej = 0;               // rejected CG steps of the last conjugate_gradient call
    DevMem x, hx, hp, g, pg, p, r, q, f, tmp, io;
    DevMem coef, part, dots, G, T, Z, w, e, info, scal, flags;
};
static KpointState *const g_kp = new KpointState();

constexpr int kZChunks = 64;     // partial sums per band and inner product
constexpr int kZMaxPairs = 6;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256)
k_z_split(const double2 *__restrict__ in, double *__restrict__ out, long ML, int nb)
{
    const long n = ML * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ML, i = t - b * ML;
        const double2 v = in[t];
        out[(2 * b) * ML + i] = v.x;
        out[(2 * b + 1) * ML + i] = v.y;
    }
}

__global__ void __launch_bounds__(256)
k_z_merge(const double *__restrict__ in, double2 *__restrict__ out, long ML, int nb)
{
    const long n = ML * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ML, i = t - b * ML;
        out[t] = make_double2(in[(2 * b) * ML + i], in[(2 * b + 1) * ML + i]);
    }
}

__device__ __forceinline__ int zwrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// out += |k|^2/2 u - i sum_{d,m} coef_nabk(d,m) (u(x + m e_d) - u(x - m e_d)):
// the part of zcoef_kin beyond the real Laplacian, on (re, im) row pairs
__global__ void __launch_bounds__(256)
k_z_kterm(const double *__restrict__ in, double *__restrict__ out, int N1, int N2, int N3, int Md,
          ZNab nab, double ck2, int nb)
{
    const long ML = (long)N1 * N2 * N3;
    const long n = ML * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ML, i = t - b * ML;
        const int i1 = (int)(i % N1);
        const int i2 = (int)((i / N1) % N2);
        const int i3 = (int)(i / ((long)N1 * N2));
        const double *ur = in + (2 * b) * ML, *ui = ur + ML;
        double sr = 0.0, si = 0.0;
        for (int m = 1; m <= Md; ++m) {
            const long xp = zwrap(i1 + m, N1) + (long)N1 * (i2 + (long)N2 * i3);
            const long xm = zwrap(i1 - m, N1) + (long)N1 * (i2 + (long)N2 * i3);
            const long yp = i1 + (long)N1 * (zwrap(i2 + m, N2) + (long)N2 * i3);
            const long ym = i1 + (long)N1 * (zwrap(i2 - m, N2) + (long)N2 * i3);
            const long zp = i1 + (long)N1 * (i2 + (long)N2 * zwrap(i3 + m, N3));
            const long zm = i1 + (long)N1 * (i2 + (long)N2 * zwrap(i3 - m, N3));
            const double a = nab.c[0][m - 1], bb = nab.c[1][m - 1], c = nab.c[2][m - 1];
            sr += a * (ui[xp] - ui[xm]) + bb * (ui[yp] - ui[ym]) + c * (ui[zp] - ui[zm]);
            si -= a * (ur[xp] - ur[xm]) + bb * (ur[yp] - ur[ym]) + c * (ur[zp] - ur[zm]);
        }
        out[(2 * b) * ML + i] += ck2 * ur[i] + sr;
        out[(2 * b + 1) * ML + i] += ck2 * ui[i] + si;
    }
}

// coef[b][l] = iuV dV sum_j conj(uVk_j) u(JJP_j): one 64-lane wave per
// (projector, band), a fixed-order lane reduction
__global__ void __launch_bounds__(64)
k_z_project(const double *__restrict__ in, long ML, int nb, int nz, const int *__restrict__ lptr,
            const int *__restrict__ lidx, const double *__restrict__ lval,
            const double *__restrict__ scale, double *__restrict__ coef)
{
    const int l = blockIdx.x / nb;
    const int b = blockIdx.x - l * nb;
    if (l >= nz) return;
    const double *ur = in + (2L * b) * ML, *ui = ur + ML;
    double cr = 0.0, ci = 0.0;
    for (int j = lptr[l] + threadIdx.x; j < lptr[l + 1]; j += 64) {
        const double vr = lval[2 * j], vi = lval[2 * j + 1];
        const int gi = lidx[j];
        const double ar = ur[gi], ai = ui[gi];
        cr += vr * ar + vi * ai;
        ci += vr * ai - vi * ar;
    }
    for (int off = 32; off > 0; off >>= 1) {
        cr += __shfl_down(cr, off, 64);
        ci += __shfl_down(ci, off, 64);
    }
    if (threadIdx.x == 0) {
        coef[2 * ((long)b * nz + l)] = scale[l] * cr;
        coef[2 * ((long)b * nz + l) + 1] = scale[l] * ci;
    }
}

// out(p) += sum over the projectors covering point p of uVk(p) coef (points
// fastest: consecutive threads, consecutive covered points)
__global__ void __launch_bounds__(256)
k_z_apply(double *__restrict__ out, long ML, int nb, int nz, int ncov,
          const int *__restrict__ cptr, const int *__restrict__ cpt, const int *__restrict__ clma,
          const double *__restrict__ cval, const double *__restrict__ coef)
{
    const long n = (long)ncov * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ncov;
        const int p = (int)(t - b * ncov);
        double hr = 0.0, hi = 0.0;
        for (int e = cptr[p]; e < cptr[p + 1]; ++e) {
            const double vr = cval[2 * e], vi = cval[2 * e + 1];
            const long ci = 2 * (b * nz + clma[e]);
            const double cr = coef[ci], cim = coef[ci + 1];
            hr += vr * cr - vi * cim;
            hi += vr * cim + vi * cr;
        }
        const long g = cpt[p];
        out[(2 * b) * ML + g] += hr;
        out[(2 * b + 1) * ML + g] += hi;
    }
}

struct ZPairs {
    const double *u[kZMaxPairs];
    const double *v[kZMaxPairs];
};

// partial sums of sum conj(u) v = (ur vr + ui vi) + i (ur vi - ui vr) per
// band, pair and chunk of the grid; grid = nb * kZChunks blocks of 256
__global__ void __launch_bounds__(256)
k_z_dots(ZPairs P, int npair, long ML, int nb, double *__restrict__ part)
{
    const int b = blockIdx.x / kZChunks;
    const int ch = blockIdx.x - b * kZChunks;
    const long len = (ML + kZChunks - 1) / kZChunks;
    const long i0 = ch * len, i1 = min(ML, i0 + len);
    __shared__ double red[256];
    for (int k = 0; k < npair; ++k) {
        const double *ur = P.u[k] + (2L * b) * ML, *ui = ur + ML;
        const double *vr = P.v[k] + (2L * b) * ML, *vi = vr + ML;
        double sr = 0.0, si = 0.0;
        for (long i = i0 + threadIdx.x; i < i1; i += 256) {
            sr += ur[i] * vr[i] + ui[i] * vi[i];
            si += ur[i] * vi[i] - ui[i] * vr[i];
        }
        for (int part_i = 0; part_i < 2; ++part_i) {
            red[threadIdx.x] = part_i == 0 ? sr : si;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
                __syncthreads();
            }
            if (threadIdx.x == 0) part[(((long)b * npair + k) * 2 + part_i) * kZChunks + ch] = red[0];
            __syncthreads();
        }
    }
}

__global__ void k_z_dots_final(const double *__restrict__ part, int n, double scale,
                               double *__restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double s = 0.0;
    for (int c = 0; c < kZChunks; ++c) s += part[(long)t * kZChunks + c];
    out[t] = s * scale;
}

// per-band complex scalars, sc[b*8 + ...]: the kernels below read the slots
// they need; flag[b] = 0 leaves band b untouched
// y = alpha y + beta x with complex alpha, beta (slots ia, ib)
__global__ void __launch_bounds__(256)
k_z_axpby(double *__restrict__ y, const double *__restrict__ x, long ML, int nb,
          const double *__restrict__ sc, int ia, int ib, const int *__restrict__ flag)
{
    const long n = ML * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ML, i = t - b * ML;
        if (flag && !flag[b]) continue;
        const double ar = sc[b * 8 + ia], ai = sc[b * 8 + ia + 1];
        const double br = sc[b * 8 + ib], bi = sc[b * 8 + ib + 1];
        double *yr = y + (2 * b) * ML + i, *yi = yr + ML;
        const double *xr = x + (2 * b) * ML + i, *xi = xr + ML;
        const double r = ar * *yr - ai * *yi + br * *xr - bi * *xi;
        const double im = ar * *yi + ai * *yr + br * *xi + bi * *xr;
        *yr = r;
        *yi = im;
    }
}

// g = -2 (hx - E (u1 x + u2 p)): E real (slot 6), u1, u2 complex (slots 0, 2);
// with p = null the initial gradient -2 (hx - E x)
__global__ void __launch_bounds__(256)
k_z_grad(double *__restrict__ g, const double *__restrict__ hx, const double *__restrict__ x,
         const double *__restrict__ p, long ML, int nb, const double *__restrict__ sc,
         const int *__restrict__ flag)
{
    const long n = ML * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long b = t / ML, i = t - b * ML;
        if (flag && !flag[b]) continue;
        const long o = (2 * b) * ML + i;
        const double E = sc[b * 8 + 6];
        double tr = x[o], ti = x[o + ML];
        if (p) {
            const double u1r = sc[b * 8], u1i = sc[b * 8 + 1];
            const double u2r = sc[b * 8 + 2], u2i = sc[b * 8 + 3];
            const double xr = tr, xi = ti;
            tr = u1r * xr - u1i * xi + u2r * p[o] - u2i * p[o + ML];
            ti = u1r * xi + u1i * xr + u2r * p[o + ML] + u2i * p[o];
        }
        g[o] = -2.0 * (hx[o] - E * tr);
        g[o + ML] = -2.0 * (hx[o + ML] - E * ti);
    }
}

// f = A g with the real 7-point operator of precond_cg_mat2 (E0 = 0), on the
// rows of the flagged bands
__global__ void __launch_bounds__(256)
k_z_pc7(const double *__restrict__ in, double *__restrict__ out, int N1, int N2, int N3,
        double c, double c1, double c2, double c3, int nrow, const int *__restrict__ flag)
{
    const long ML = (long)N1 * N2 * N3;
    const long n = ML * nrow;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long r = t / ML, i = t - r * ML;
        if (flag && !flag[r / 2]) continue;
        const int i1 = (int)(i % N1), i2 = (int)((i / N1) % N2), i3 = (int)(i / ((long)N1 * N2));
        const double *s = in + r * ML;
        const double v = c * s[i]
            + c1 * (s[zwrap(i1 + 1, N1) + (long)N1 * (i2 + (long)N2 * i3)] +
                    s[zwrap(i1 - 1, N1) + (long)N1 * (i2 + (long)N2 * i3)])
            + c2 * (s[i1 + (long)N1 * (zwrap(i2 + 1, N2) + (long)N2 * i3)] +
                    s[i1 + (long)N1 * (zwrap(i2 - 1, N2) + (long)N2 * i3)])
            + c3 * (s[i1 + (long)N1 * (i2 + (long)N2 * zwrap(i3 + 1, N3))] +
                    s[i1 + (long)N1 * (i2 + (long)N2 * zwrap(i3 - 1, N3))]);
        out[r * ML + i] = v;
    }
}

// complex nb x nb (column-major) from the real row Gram matrix G (2nb x 2nb,
// column-major): C_mn = s [G(2m,2n) + G(2m+1,2n+1) + i (G(2m,2n+1) - G(2m+1,2n))]
// = s sum conj(x_m) y_n; herm = 1 stores (C + C^H)/2
__global__ void __launch_bounds__(256)
k_z_from_gram(const double *__restrict__ G, int nb, double s, int herm, double2 *__restrict__ C)
{
    const long n = (long)nb * nb;
    const long ld = 2L * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long m = t % nb, k = t / nb;
        double re = G[(2 * m) + ld * (2 * k)] + G[(2 * m + 1) + ld * (2 * k + 1)];
        double im = G[(2 * m) + ld * (2 * k + 1)] - G[(2 * m + 1) + ld * (2 * k)];
        if (herm) {
            const double re2 = G[(2 * k) + ld * (2 * m)] + G[(2 * k + 1) + ld * (2 * m + 1)];
            const double im2 = G[(2 * k) + ld * (2 * m + 1)] - G[(2 * k + 1) + ld * (2 * m)];
            re = 0.5 * (re + re2);
            im = 0.5 * (im - im2);
        }
        C[t] = make_double2(s * re, s * im);
    }
}

// T (2nb x 2nb, column-major): the real form of right-multiplying the complex
// rows by Z (q_n = sum_m x_m Z_mn); upper = 1 takes Z as upper triangular
__global__ void __launch_bounds__(256)
k_z_block(const double2 *__restrict__ Z, int nb, int upper, double *__restrict__ T)
{
    const long n = (long)nb * nb;
    const long ld = 2L * nb;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const long m = t % nb, k = t / nb;
        double2 z = Z[t];
        if (upper && m > k) z = make_double2(0.0, 0.0);
        T[(2 * m) + ld * (2 * k)] = z.x;
        T[(2 * m + 1) + ld * (2 * k)] = -z.y;
        T[(2 * m) + ld * (2 * k + 1)] = z.y;
        T[(2 * m + 1) + ld * (2 * k + 1)] = z.x;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static unsigned grid_for(long n)
{
    return (unsigned)std::min<long>((n + 255) / 256, 8192);
}

static int z_buffers(int MB)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const size_t row = (size_t)c.g.ML * 2 * MB * sizeof(double);
    for (DevMem *m : {&K.x, &K.hx, &K.hp, &K.g, &K.pg, &K.p, &K.r, &K.q, &K.f, &K.tmp, &K.io})
        RS_TRY(m->ensure(row));
    RS_TRY(K.coef.ensure(sizeof(double) * 2 * (size_t)MB * std::max(1, K.nz)));
    RS_TRY(K.part.ensure(sizeof(double) * (size_t)MB * kZMaxPairs * 2 * kZChunks));
    RS_TRY(K.dots.ensure(sizeof(double) * (size_t)MB * kZMaxPairs * 2));
    RS_TRY(K.G.ensure(sizeof(double) * 4 * (size_t)MB * MB));
    RS_TRY(K.T.ensure(sizeof(double) * 4 * (size_t)MB * MB));
    RS_TRY(K.Z.ensure(sizeof(double2) * (size_t)MB * MB));
    RS_TRY(K.w.ensure(sizeof(double) * MB));
    RS_TRY(K.e.ensure(sizeof(double) * MB));
    RS_TRY(K.info.ensure(sizeof(int) * 2));
    RS_TRY(K.scal.ensure(sizeof(double) * 8 * MB));
    RS_TRY(K.flags.ensure(sizeof(int) * MB));
    K.MB = MB;
    return kOK;
}

// H u for nb complex bands in split layout
static int zhpsi(const double *in, double *out, int nb)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const GridDesc &g = c.g;
    RS_TRY(stencil_kinloc<double>(in, out, 2 * nb));
    hipLaunchKernelGGL(k_z_kterm, dim3(grid_for(g.ML * nb)), dim3(256), 0, c.stream, in, out, g.N1,
                       g.N2, g.N3, g.Md, K.nab, K.ck2, nb);
    RS_HIP(hipGetLastError());
    if (K.nz > 0 && K.ncov > 0) {
        hipLaunchKernelGGL(k_z_project, dim3((unsigned)((long)K.nz * nb)), dim3(64), 0, c.stream, in,
                           g.ML, nb, K.nz, K.lptr.as<int>(), K.lidx.as<int>(), K.lval.as<double>(),
                           K.scale.as<double>(), K.coef.as<double>());
        hipLaunchKernelGGL(k_z_apply, dim3(grid_for((long)K.ncov * nb)), dim3(256), 0, c.stream,
                           out, g.ML, nb, K.nz, K.ncov, K.cptr.as<int>(), K.cpt.as<int>(),
                           K.clma.as<int>(), K.cval.as<double>(), K.coef.as<double>());
        RS_HIP(hipGetLastError());
    }
    return kOK;
}

// npair inner products dV sum conj(u_k) v_k per band, to the host: out[(b*npair + k)*2 + re/im]
static int zdots(const ZPairs &P, int npair, int nb, double scale, std::vector<double> &out)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    hipLaunchKernelGGL(k_z_dots, dim3((unsigned)(nb * kZChunks)), dim3(256), 0, c.stream, P, npair,
                       c.g.ML, nb, K.part.as<double>());
    const int n = nb * npair * 2;
    hipLaunchKernelGGL(k_z_dots_final, dim3((n + 255) / 256), dim3(256), 0, c.stream,
                       K.part.as<double>(), n, scale, K.dots.as<double>());
    RS_HIP(hipGetLastError());
    out.resize(n);
    RS_HIP(hipMemcpyAsync(out.data(), K.dots.as<double>(), sizeof(double) * n,
                          hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

static int upload_scal(const std::vector<double> &sc, const std::vector<int> &flag)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    RS_HIP(hipMemcpyAsync(K.scal.as<double>(), sc.data(), sizeof(double) * sc.size(),
                          hipMemcpyHostToDevice, c.stream));
    RS_HIP(hipMemcpyAsync(K.flags.as<int>(), flag.data(), sizeof(int) * flag.size(),
                          hipMemcpyHostToDevice, c.stream));
    // the host vectors are reused right after: wait for the copies
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

// zhegv(1,'V','U') of the 2x2 Hermitian pair (A, B): ascending w, z^H B z = 1
static void hegv2(const zc A[2][2], const zc B[2][2], double w[2], zc Z[2][2])
{
    // B = L L^H
    const double l11 = std::sqrt(B[0][0].real());
    const zc l21 = B[1][0] / l11;
    const double l22 = std::sqrt(B[1][1].real() - std::norm(l21));
    // C = L^-1 A L^-H (Hermitian)
    const zc i11 = 1.0 / l11, i22 = 1.0 / l22, i21 = -l21 / (l11 * l22);
    auto Li = [&](int r, int s) -> zc { return r == s ? (r == 0 ? i11 : i22) : (r == 1 && s == 0 ? i21 : zc(0.0)); };
    zc C[2][2];
    for (int r = 0; r < 2; ++r)
        for (int s = 0; s < 2; ++s) {
            zc t = 0.0;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) t += Li(r, a) * A[a][b] * std::conj(Li(s, b));
            C[r][s] = t;
        }
    const double c11 = C[0][0].real(), c22 = C[1][1].real();
    const zc c12 = C[0][1];
    const double h = 0.5 * (c11 - c22);
    const double rad = std::sqrt(h * h + std::norm(c12));
    const double m = 0.5 * (c11 + c22);
    w[0] = m - rad;
    w[1] = m + rad;
    zc Y[2][2];
    if (std::abs(c12) == 0.0) {
        const bool keep = c11 <= c22;
        Y[0][0] = keep ? 1.0 : 0.0; Y[1][0] = keep ? 0.0 : 1.0;
        Y[0][1] = keep ? 0.0 : 1.0; Y[1][1] = keep ? 1.0 : 0.0;
    } else {
        for (int k = 0; k < 2; ++k) {
            // (C - w I) y = 0: y = (c12, w - c11) normalised
            zc y0 = c12, y1 = w[k] - c11;
            if (std::abs(y0) + std::abs(y1) < 1e-300) { y0 = 1.0; y1 = 0.0; }
            // the better conditioned of the two row choices
            const zc z0 = w[k] - c22, z1 = std::conj(c12);
            if (std::abs(z0) + std::abs(z1) > std::abs(y0) + std::abs(y1)) { y0 = z0; y1 = z1; }
            const double nrm = std::sqrt(std::norm(y0) + std::norm(y1));
            Y[0][k] = y0 / nrm;
            Y[1][k] = y1 / nrm;
        }
    }
    // z = L^-H y
    for (int k = 0; k < 2; ++k) {
        const zc y0 = Y[0][k], y1 = Y[1][k];
        Z[1][k] = std::conj(i22) * y1;
        Z[0][k] = std::conj(i11) * y0 + std::conj(i21) * y1;
    }
}

// precond_hprsdft on the flagged bands: pg = g + sum a_k p_k from x0 = g
static int zprecond(int nb, const std::vector<int> &flag)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const GridDesc &g = c.g;
    const long ML = g.ML;
    if (c.mloop <= 0) return kOK;
    const double H1 = g.h[0], H2 = g.h[1], H3 = g.h[2];
    const double cc = 1.0 / (H1 * H1) + 1.0 / (H2 * H2) + 1.0 / (H3 * H3);
    const double c1 = -0.5 / (H1 * H1), c2 = -0.5 / (H2 * H2), c3 = -0.5 / (H3 * H3);
    double *gk = K.g.as<double>(), *pg = K.pg.as<double>(), *rk = K.r.as<double>(),
           *pk = K.q.as<double>(), *f = K.f.as<double>();
    std::vector<int> fl(flag);
    std::vector<double> sc(8 * (size_t)nb, 0.0), d;
    RS_TRY(upload_scal(sc, fl));
    hipLaunchKernelGGL(k_z_pc7, dim3(grid_for(2 * ML * nb)), dim3(256), 0, c.stream, gk, f, g.N1,
                       g.N2, g.N3, cc, c1, c2, c3, 2 * nb, K.flags.as<int>());
    // r = g - f (slots 0 and 2: alpha, beta), p = r
    RS_HIP(hipMemcpyAsync(rk, gk, sizeof(double) * 2 * ML * nb, hipMemcpyDeviceToDevice, c.stream));
    for (int b = 0; b < nb; ++b) { sc[8 * b] = 1.0; sc[8 * b + 2] = -1.0; }
    RS_TRY(upload_scal(sc, fl));
    hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, rk, f, ML, nb,
                       K.scal.as<double>(), 0, 2, K.flags.as<int>());
    RS_HIP(hipMemcpyAsync(pk, rk, sizeof(double) * 2 * ML * nb, hipMemcpyDeviceToDevice, c.stream));
    ZPairs P{};
    P.u[0] = rk; P.v[0] = rk;
    RS_TRY(zdots(P, 1, nb, 1.0, d));
    std::vector<double> rr1(nb), rr0(nb);
    for (int b = 0; b < nb; ++b) {
        rr1[b] = d[2 * b];
        if (fl[b] && rr1[b] < 1.0e-24) fl[b] = 0;       // the early return: pg stays g
    }
    for (int it = 1; it <= c.mloop; ++it) {
        RS_TRY(upload_scal(sc, fl));
        hipLaunchKernelGGL(k_z_pc7, dim3(grid_for(2 * ML * nb)), dim3(256), 0, c.stream, pk, f,
                           g.N1, g.N2, g.N3, cc, c1, c2, c3, 2 * nb, K.flags.as<int>());
        P.u[0] = pk; P.v[0] = f;
        RS_TRY(zdots(P, 1, nb, 1.0, d));
        std::vector<double> a(nb, 0.0);
        for (int b = 0; b < nb; ++b) {
            rr0[b] = rr1[b];
            a[b] = fl[b] ? rr0[b] / d[2 * b] : 0.0;
        }
        for (int b = 0; b < nb; ++b) {
            sc[8 * b] = 1.0; sc[8 * b + 1] = 0.0;
            sc[8 * b + 2] = a[b]; sc[8 * b + 3] = 0.0;
        }
        RS_TRY(upload_scal(sc, fl));
        hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, pg, pk, ML,
                           nb, K.scal.as<double>(), 0, 2, K.flags.as<int>());
        if (it == c.mloop) break;
        for (int b = 0; b < nb; ++b) sc[8 * b + 2] = -a[b];
        RS_TRY(upload_scal(sc, fl));
        hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, rk, f, ML,
                           nb, K.scal.as<double>(), 0, 2, K.flags.as<int>());
        P.u[0] = rk; P.v[0] = rk;
        RS_TRY(zdots(P, 1, nb, 1.0, d));
        for (int b = 0; b < nb; ++b) {
            rr1[b] = d[2 * b];
            const double beta = fl[b] ? rr1[b] / rr0[b] : 0.0;
            sc[8 * b] = beta; sc[8 * b + 1] = 0.0;
            sc[8 * b + 2] = 1.0; sc[8 * b + 3] = 0.0;
        }
        RS_TRY(upload_scal(sc, fl));
        hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, pk, rk, ML,
                           nb, K.scal.as<double>(), 0, 2, K.flags.as<int>());
        RS_HIP(hipGetLastError());
    }
    return kOK;
}

// conjugate_gradient_1, complex branch, all bands at once (each band's steps
// are independent; MB_d = 1 semantics)
static int zcg(int nb, std::vector<double> &esp, std::vector<double> &res)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const long ML = c.g.ML;
    const double dV = c.g.dV, c1 = 2.0;
    double *x = K.x.as<double>(), *hx = K.hx.as<double>(), *hp = K.hp.as<double>(),
           *gk = K.g.as<double>(), *pg = K.pg.as<double>(), *pk = K.p.as<double>();
    const size_t bytes = sizeof(double) * 2 * ML * nb;
    std::vector<double> sc(8 * (size_t)nb, 0.0), d;
    std::vector<int> all(nb, 1);
    RS_TRY(zhpsi(x, hx, nb));
    ZPairs P{};
    P.u[0] = x; P.v[0] = hx;
    RS_TRY(zdots(P, 1, nb, dV, d));
    std::vector<double> E(nb), E1(nb, 1.0e10), rb(nb), gkgk(nb, 0.0);
    for (int b = 0; b < nb; ++b) { E[b] = d[2 * b]; sc[8 * b + 6] = E[b]; }
    RS_TRY(upload_scal(sc, all));
    hipLaunchKernelGGL(k_z_grad, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, gk, hx, x,
                       (const double *)nullptr, ML, nb, K.scal.as<double>(), K.flags.as<int>());
    P.u[0] = gk; P.v[0] = gk;
    RS_TRY(zdots(P, 1, nb, dV, d));
    for (int b = 0; b < nb; ++b) rb[b] = d[2 * b];
    std::vector<int> run(nb, 1);
    esp.assign(nb, 0.0);
    res.assign(nb, 0.0);
    int nrej = 0;
    for (int icg = 1; icg <= c.ncg + 1; ++icg) {
        int nrun = 0;
        for (int b = 0; b < nb; ++b) {
            if (!run[b]) continue;
            res[b] = rb[b] / (c1 * c1);
            if (rb[b] < 0.0 || std::fabs(E[b] - E1[b]) < 1.0e-15 || icg == c.ncg + 1) run[b] = 0;
            else ++nrun;
        }
        if (nrun == 0) break;
        RS_HIP(hipMemcpyAsync(pg, gk, bytes, hipMemcpyDeviceToDevice, c.stream));
        RS_TRY(zprecond(nb, run));
        P.u[0] = pg; P.v[0] = gk;
        RS_TRY(zdots(P, 1, nb, dV, d));
        for (int b = 0; b < nb; ++b) {
            const double r = d[2 * b];
            // p = Pg + (rb / gkgk) p after the first step
            sc[8 * b] = icg == 1 ? 0.0 : r / gkgk[b]; sc[8 * b + 1] = 0.0;
            sc[8 * b + 2] = 1.0; sc[8 * b + 3] = 0.0;
            if (run[b]) gkgk[b] = r;
        }
        if (icg == 1) {
            RS_HIP(hipMemcpyAsync(pk, pg, bytes, hipMemcpyDeviceToDevice, c.stream));
        } else {
            RS_TRY(upload_scal(sc, run));
            hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, pk, pg,
                               ML, nb, K.scal.as<double>(), 0, 2, K.flags.as<int>());
        }
        RS_TRY(zhpsi(pk, hp, nb));
        P.u[0] = x;  P.v[0] = x;
        P.u[1] = pk; P.v[1] = x;
        P.u[2] = pk; P.v[2] = pk;
        P.u[3] = x;  P.v[3] = hx;
        P.u[4] = pk; P.v[4] = hx;
        P.u[5] = pk; P.v[5] = hp;
        RS_TRY(zdots(P, 6, nb, dV, d));
        for (int b = 0; b < nb; ++b) {
            if (!run[b]) continue;
            const double *v = d.data() + 12 * (size_t)b;
            const zc v1(v[0], 0.0), v2(v[2], v[3]), v3(v[4], 0.0);
            const zc v4(v[6], 0.0), v5(v[8], v[9]), v6(v[10], 0.0);
            const zc Bm[2][2] = {{v1, std::conj(v2)}, {v2, v3}};
            const zc Am[2][2] = {{v4, std::conj(v5)}, {v5, v6}};
            double W[2];
            zc Z[2][2];
            hegv2(Am, Bm, W, Z);
            int k = 0;
            if (std::fabs(W[0] - E[b]) > 1.0e-1 && std::fabs(W[1] - E[b]) <= 1.0e-1) k = 1;
            zc u1 = Z[0][k], u2 = Z[1][k];
            const double r = std::abs(u1);
            const zc ph(u1.real() / r, -u1.imag() / r);
            u1 *= ph;
            u2 *= ph;
            E1[b] = E[b];
            E[b] = W[k];
            sc[8 * b] = u1.real(); sc[8 * b + 1] = u1.imag();
            sc[8 * b + 2] = u2.real(); sc[8 * b + 3] = u2.imag();
            sc[8 * b + 6] = W[k];
        }
        RS_TRY(upload_scal(sc, run));
        // hx = u1 hx + u2 hp; g = -2 (hx - w (u1 x + u2 p)); rb = |g|^2
        hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, hx, hp, ML,
                           nb, K.scal.as<double>(), 0, 2, K.flags.as<int>());
        hipLaunchKernelGGL(k_z_grad, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, gk, hx, x,
                           (const double *)pk, ML, nb, K.scal.as<double>(), K.flags.as<int>());
        P.u[0] = gk; P.v[0] = gk;
        RS_TRY(zdots(P, 1, nb, dV, d));
        std::vector<int> upd(nb, 0);
        for (int b = 0; b < nb; ++b) {
            if (!run[b]) continue;
            rb[b] = d[2 * b];
            if (rb[b] / res[b] > 1.0e8) { E[b] = E1[b]; ++nrej; continue; }
            upd[b] = 1;
        }
        RS_TRY(upload_scal(sc, upd));
        hipLaunchKernelGGL(k_z_axpby, dim3(grid_for(ML * nb)), dim3(256), 0, c.stream, x, pk, ML, nb,
                           K.scal.as<double>(), 0, 2, K.flags.as<int>());
        RS_HIP(hipGetLastError());
    }
    for (int b = 0; b < nb; ++b) esp[b] = E[b];
    K.nrej = nrej;
    return kOK;
}

// x <- x Z on the row matrix (upper = 1: Z upper triangular)
static int zrotate(double *x, int nb, const double2 *Z, int upper)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const long ML = c.g.ML;
    const int n2 = 2 * nb;
    hipLaunchKernelGGL(k_z_block, dim3(grid_for((long)nb * nb)), dim3(256), 0, c.stream, Z, nb, upper,
                       K.T.as<double>());
    RS_HIP(hipGetLastError());
    const double one = 1.0, zero = 0.0;
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, (int)ML, n2, n2,
                          &one, x, (int)ML, K.T.as<double>(), n2, &zero, K.tmp.as<double>(), (int)ML));
    RS_HIP(hipMemcpyAsync(x, K.tmp.as<double>(), sizeof(double) * ML * n2, hipMemcpyDeviceToDevice,
                          c.stream));
    return kOK;
}

// C = s X^H Y as a complex nb x nb matrix in K.Z (herm: its Hermitian part)
static int zgram(const double *X, const double *Y, int nb, double s, int herm)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const long ML = c.g.ML;
    const int n2 = 2 * nb;
    const double one = 1.0, zero = 0.0;
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n2, n2,
                          (int)ML, &one, X, (int)ML, Y, (int)ML, &zero, K.G.as<double>(), n2));
    hipLaunchKernelGGL(k_z_from_gram, dim3(grid_for((long)nb * nb)), dim3(256), 0, c.stream,
                       K.G.as<double>(), nb, s, herm, K.Z.as<double2>());
    RS_HIP(hipGetLastError());
    return kOK;
}

static int check_info()
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    int info = 0;
    RS_HIP(hipMemcpyAsync(&info, K.info.as<int>(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    if (info != 0) return set_error(kErrSolver, "complex factorisation failed, info %d", info);
    return kOK;
}

// Gram-Schmidt: Q = X U^-1 with dV X^H X = U^H U (U upper, positive real diagonal)
static int zgs(int nb)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    RS_TRY(zgram(K.x.as<double>(), K.x.as<double>(), nb, c.g.dV, 1));
    auto *Z = reinterpret_cast<rocblas_double_complex *>(K.Z.as<double2>());
    RS_BLAS(rocsolver_zpotrf(c.blas, rocblas_fill_upper, nb, Z, nb, K.info.as<int>()));
    RS_TRY(check_info());
    RS_BLAS(rocsolver_ztrtri(c.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, nb, Z, nb,
                             K.info.as<int>()));
    RS_TRY(check_info());
    return zrotate(K.x.as<double>(), nb, K.Z.as<double2>(), 1);
}

// subspace_diag_la: Hsub = dV (X^H HX + (HX)^H X)/2, zheevd ascending, X <- X V
static int zsd(int nb, std::vector<double> &esp)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    RS_TRY(zhpsi(K.x.as<double>(), K.hx.as<double>(), nb));
    RS_TRY(zgram(K.x.as<double>(), K.hx.as<double>(), nb, c.g.dV, 1));
    auto *Z = reinterpret_cast<rocblas_double_complex *>(K.Z.as<double2>());
    RS_BLAS(rocsolver_zheevd(c.blas, rocblas_evect_original, rocblas_fill_upper, nb, Z, nb,
                             K.w.as<double>(), K.e.as<double>(), K.info.as<int>()));
    RS_TRY(check_info());
    esp.resize(nb);
    RS_HIP(hipMemcpyAsync(esp.data(), K.w.as<double>(), sizeof(double) * nb, hipMemcpyDeviceToHost,
                          c.stream));
    RS_TRY(zrotate(K.x.as<double>(), nb, K.Z.as<double2>(), 0));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

static int kp_ready(int ML0, int MB)
{
    Context &c = ctx();
    if (!c.init) return set_error(kErrNotInit, "rsdft_gpu_init has not been called");
    if (!g_kp->set) return set_error(kErrState, "rsdft_gpu_set_kpoint has not been called");
    if (ML0 != c.g.ML) return set_error(kErrArg, "ML0=%d, grid has %ld points", ML0, c.g.ML);
    if (MB <= 0) return set_error(kErrArg, "MB=%d", MB);
    return z_buffers(MB);
}

static int load_x(const double *unk, int MB)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const long ML = c.g.ML;
    RS_HIP(hipMemcpyAsync(K.io.as<double>(), unk, sizeof(double) * 2 * ML * MB, hipMemcpyHostToDevice,
                          c.stream));
    hipLaunchKernelGGL(k_z_split, dim3(grid_for(ML * MB)), dim3(256), 0, c.stream,
                       K.io.as<double2>(), K.x.as<double>(), ML, MB);
    RS_HIP(hipGetLastError());
    return kOK;
}

static int store(const double *src, double *dst, int MB)
{
    KpointState &K = *g_kp;
    Context &c = ctx();
    const long ML = c.g.ML;
    hipLaunchKernelGGL(k_z_merge, dim3(grid_for(ML * MB)), dim3(256), 0, c.stream, src,
                       K.io.as<double2>(), ML, MB);
    RS_HIP(hipGetLastError());
    RS_HIP(hipMemcpyAsync(dst, K.io.as<double>(), sizeof(double) * 2 * ML * MB, hipMemcpyDeviceToHost,
                          c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

void kpoint_release()
{
    KpointState &K = *g_kp;
    for (DevMem *m : {&K.lptr, &K.lidx, &K.lval, &K.scale, &K.cptr, &K.cpt, &K.clma, &K.cval, &K.x,
                      &K.hx, &K.hp, &K.g, &K.pg, &K.p, &K.r, &K.q, &K.f, &K.tmp, &K.io, &K.coef,
                      &K.part, &K.dots, &K.G, &K.T, &K.Z, &K.w, &K.e, &K.info, &K.scal, &K.flags})
        m->release();
    K.set = false;
    K.nz = K.ncov = K.MB = 0;
}

}  // namespace rsdft

using namespace rsdft;

extern "C" {

RSDFT_API int rsdft_gpu_set_kpoint(double const_k2, const double *coef_nabk, int nzlma, int maxmjj,
                                   const int *MJJ, const int *JJP, const double *uVk,
                                   const int *iuV)
{
    Context &c = ctx();
    if (!c.init) return set_error(kErrNotInit, "rsdft_gpu_init has not been called");
    KpointState &K = *g_kp;
    const int Md = c.g.Md;
    const long ML = c.g.ML;
    K.ck2 = const_k2;
    for (int d = 0; d < 3; ++d)
        for (int m = 0; m < kMaxMd; ++m) K.nab.c[d][m] = m < Md ? coef_nabk[d + 3 * m] : 0.0;
    // per projector: (point, uVk) sorted by point, repeated points summed
    using Entry = std::pair<int, zc>;
    std::vector<std::vector<Entry>> pts(std::max(nzlma, 0));
    for (int l = 0; l < nzlma; ++l) {
        if (MJJ[l] < 0 || MJJ[l] > maxmjj) return set_error(kErrArg, "MJJ(%d)=%d", l + 1, MJJ[l]);
        auto &P = pts[l];
        for (int j = 0; j < MJJ[l]; ++j) {
            const long o = (long)l * maxmjj + j;
            const int gi = JJP[o] - 1;
            if (gi < 0 || gi >= ML)
                return set_error(kErrArg, "JJP(%d,%d)=%d out of grid", j + 1, l + 1, gi + 1);
            P.push_back({gi, zc(uVk[2 * o], uVk[2 * o + 1])});
        }
        std::stable_sort(P.begin(), P.end(),
                         [](const Entry &a, const Entry &b) { return a.first < b.first; });
        size_t w = 0;
        for (size_t k = 0; k < P.size(); ++k) {
            if (w > 0 && P[w - 1].first == P[k].first) P[w - 1].second += P[k].second;
            else P[w++] = P[k];
        }
        P.resize(w);
    }
    std::vector<int> lptr(nzlma + 1, 0), lidx;
    std::vector<double> lval, scale(std::max(nzlma, 1));
    for (int l = 0; l < nzlma; ++l) {
        for (auto &e : pts[l]) {
            lidx.push_back(e.first);
            lval.push_back(e.second.real());
            lval.push_back(e.second.imag());
        }
        lptr[l + 1] = (int)lidx.size();
        scale[l] = iuV[l] * c.g.dV;
    }
    // point-major transpose: covered points ascending, their projectors ascending
    std::vector<std::vector<std::pair<int, zc>>> byp;
    std::vector<int> cpt;
    {
        std::vector<int> slot(ML, -1);
        for (int l = 0; l < nzlma; ++l)
            for (auto &e : pts[l])
                if (slot[e.first] < 0) { slot[e.first] = 0; }
        for (long i = 0; i < ML; ++i)
            if (slot[i] == 0) { slot[i] = (int)cpt.size(); cpt.push_back((int)i); }
        byp.resize(cpt.size());
        for (int l = 0; l < nzlma; ++l)
            for (auto &e : pts[l]) byp[slot[e.first]].push_back({l, e.second});
    }
    std::vector<int> cptr(cpt.size() + 1, 0), clma;
    std::vector<double> cval;
    for (size_t p = 0; p < cpt.size(); ++p) {
        for (auto &e : byp[p]) {
            clma.push_back(e.first);
            cval.push_back(e.second.real());
            cval.push_back(e.second.imag());
        }
        cptr[p + 1] = (int)clma.size();
    }
    auto up = [&](DevMem &m, const void *src, size_t bytes) -> int {
        RS_TRY(m.ensure(std::max<size_t>(bytes, 8)));
        if (bytes) RS_HIP(hipMemcpy(m.p, src, bytes, hipMemcpyHostToDevice));
        return kOK;
    };
    RS_TRY(up(K.lptr, lptr.data(), sizeof(int) * lptr.size()));
    RS_TRY(up(K.lidx, lidx.data(), sizeof(int) * lidx.size()));
    RS_TRY(up(K.lval, lval.data(), sizeof(double) * lval.size()));
    RS_TRY(up(K.scale, scale.data(), sizeof(double) * scale.size()));
    RS_TRY(up(K.cptr, cptr.data(), sizeof(int) * cptr.size()));
    RS_TRY(up(K.cpt, cpt.data(), sizeof(int) * cpt.size()));
    RS_TRY(up(K.clma, clma.data(), sizeof(int) * clma.size()));
    RS_TRY(up(K.cval, cval.data(), sizeof(double) * cval.size()));
    K.nz = std::max(nzlma, 0);
    K.ncov = (int)cpt.size();
    K.set = true;
    return kOK;
}

RSDFT_API int rsdft_gpu_zhamiltonian(int k, int s, const double *tpsi, double *htpsi, int n1,
                                     int n2, int ib1, int ib2)
{
    (void)s;
    if (k != 1) return set_error(kErrArg, "one k point per context (k=%d)", k);
    const int ML0 = n2 - n1 + 1, nb = ib2 - ib1 + 1;
    RS_TRY(kp_ready(ML0, nb));
    KpointState &K = *g_kp;
    RS_TRY(load_x(tpsi, nb));
    RS_TRY(zhpsi(K.x.as<double>(), K.hx.as<double>(), nb));
    return store(K.hx.as<double>(), htpsi, nb);
}

RSDFT_API int rsdft_gpu_zconjugate_gradient(double *unk, double *esp, double *res, int ML0, int MB)
{
    RS_TRY(kp_ready(ML0, MB));
    KpointState &K = *g_kp;
    RS_TRY(load_x(unk, MB));
    std::vector<double> e, r;
    RS_TRY(zcg(MB, e, r));
    std::copy(e.begin(), e.end(), esp);
    std::copy(r.begin(), r.end(), res);
    return store(K.x.as<double>(), unk, MB);
}

RSDFT_API int rsdft_gpu_zgram_schmidt(double *unk, int ML0, int MB)
{
    RS_TRY(kp_ready(ML0, MB));
    KpointState &K = *g_kp;
    RS_TRY(load_x(unk, MB));
    RS_TRY(zgs(MB));
    return store(K.x.as<double>(), unk, MB);
}

RSDFT_API int rsdft_gpu_zsubspace_diag(double *unk, double *esp, int ML0, int MB)
{
    RS_TRY(kp_ready(ML0, MB));
    KpointState &K = *g_kp;
    RS_TRY(load_x(unk, MB));
    std::vector<double> e;
    RS_TRY(zsd(MB, e));
    std::copy(e.begin(), e.end(), esp);
    return store(K.x.as<double>(), unk, MB);
}

RSDFT_API int rsdft_gpu_zsweep(double *unk, double *esp, double *res, int ML0, int MB, int nsweep)
{
    RS_TRY(kp_ready(ML0, MB));
    KpointState &K = *g_kp;
    RS_TRY(load_x(unk, MB));
    std::vector<double> e, r;
    for (int it = 0; it < nsweep; ++it) {
        RS_TRY(zcg(MB, e, r));
        RS_TRY(zgs(MB));
        RS_TRY(zsd(MB, e));
    }
    std::copy(e.begin(), e.end(), esp);
    std::copy(r.begin(), r.end(), res);
    return store(K.x.as<double>(), unk, MB);
}

}  // extern "C"
