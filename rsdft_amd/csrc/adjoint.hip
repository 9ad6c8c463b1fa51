// Adjoint (backward) kernels of the band energy E = sum_n f_n <psi_n|H|psi_n> dV
// at fixed psi:
//   dE/dVloc(r) = dV * rho(r),  rho = sum_n f_n psi_n(r)^2
//     (rho exactly as calc_density builds it, density_module.f90:129-136);
//   dE/dR_a = -force2(:,a), force2 as calc_force_ps_nloc2 builds it
//     (ps_nloc2_module.f90:1601-1773): per band and projector
//     w0 = iuV (-2 f dV^2) sum uV psi,  w = sum duVdR psi,  force2 += w0 w.
#include "common.h"

namespace rsdft {

template <class T>
__global__ void __launch_bounds__(256)
k_density(const T *__restrict__ X, long ML, int nb, const double *__restrict__ occ,
          double *__restrict__ rho)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ML;
         i += (long)gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int n = 0; n < nb; ++n) {
            const double f = occ[n];
            if (f == 0.0) continue;
            const double p = (double)X[(long)n * ML + i];
            t += f * p * p;
        }
        rho[i] = t;
    }
}

// W[b][lma][d] = sum_{j < MJJ_MAP(lma)} duVdR(d,j,lma) psi_b(JJ_MAP(j,lma))
template <class T, int NBB>
__global__ void __launch_bounds__(256)
k_dproject(const T *__restrict__ X, long ML, int nb, int nz, const int *__restrict__ dptr,
           const int *__restrict__ didx, const double *__restrict__ dval,
           const double *__restrict__ occ, double *__restrict__ W)
{
    __shared__ double red[4][3 * NBB];
    const int lma = blockIdx.x;
    const int b0 = blockIdx.y * NBB;
    double acc[3 * NBB];
#pragma unroll
    for (int k = 0; k < 3 * NBB; ++k) acc[k] = 0.0;
    bool act[NBB];
#pragma unroll
    for (int k = 0; k < NBB; ++k) act[k] = (b0 + k < nb) && occ[b0 + k] != 0.0;
    const int j1 = dptr[lma + 1];
    for (int j = dptr[lma] + threadIdx.x; j < j1; j += blockDim.x) {
        const int i = didx[j];
        const double d0 = dval[3 * (long)j], d1 = dval[3 * (long)j + 1], d2 = dval[3 * (long)j + 2];
#pragma unroll
        for (int k = 0; k < NBB; ++k) {
            if (!act[k]) continue;
            const double p = (double)X[(long)(b0 + k) * ML + i];
            acc[3 * k] += d0 * p;
            acc[3 * k + 1] += d1 * p;
            acc[3 * k + 2] += d2 * p;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3 * NBB; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3 * NBB) {
        const int k = threadIdx.x / 3;
        const int b = b0 + k;
        if (b < nb) {
            const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] +
                             red[3][threadIdx.x];
            W[((long)b * nz + lma) * 3 + (threadIdx.x - 3 * k)] = s;
        }
    }
}

// force2(d, a) = sum_b sum_{lma of a} (-2 f_b dV) C[lma][b] W[b][lma][d]
// C already carries iuV*dV, so -2 f dV^2 iuV sum uV psi = -2 f dV C.
template <class T>
__global__ void __launch_bounds__(256)
k_force_reduce(const T *__restrict__ C, const double *__restrict__ W, int nb, int nz,
               const int *__restrict__ aptr, const int *__restrict__ alma,
               const double *__restrict__ occ, double dV, double *__restrict__ force)
{
    __shared__ double red[4][3];
    const int a = blockIdx.x;
    const int l0 = aptr[a], nl = aptr[a + 1] - aptr[a];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const long npair = (long)nb * nl;
    for (long t = threadIdx.x; t < npair; t += blockDim.x) {
        const int b = (int)(t / nl);
        const int lma = alma[l0 + (int)(t - (long)b * nl)];
        const double f = occ[b];
        if (f == 0.0) continue;
        const double w0 = -2.0 * f * dV * (double)C[(long)lma * nb + b];
        const double *w = W + ((long)b * nz + lma) * 3;
        s0 += w0 * w[0];
        s1 += w0 * w[1];
        s2 += w0 * w[2];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off, 64);
        s1 += __shfl_down(s1, off, 64);
        s2 += __shfl_down(s2, off, 64);
    }
    if (lane == 0) { red[wave][0] = s0; red[wave][1] = s1; red[wave][2] = s2; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        force[3 * a + d] = red[0][d] + red[1][d] + red[2][d] + red[3][d];
    }
}

template <class T>
int density_dev(const T *X, int nb, const double *occ_dev, double *rho_dev)
{
    Context &c = ctx();
    hipLaunchKernelGGL((k_density<T>), dim3(4096), dim3(256), 0, c.stream, X, c.g.ML, nb, occ_dev,
                       rho_dev);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int force_dev(const T *X, int nb, const double *occ_dev, double *force_out)
{
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    if (nl.natom <= 0 || nl.nz <= 0)
        return set_error(kErrState, "projector derivative tables not set");
    // projection with the reference's scale iuV*dV, all bands
    RS_TRY(c.coefC.ensure((size_t)nb * nl.nz * sizeof(T)));
    RS_TRY(nl_project<T>(X, nb, c.coefC.as<T>(), nullptr));
    // derivative projections
    constexpr int NBB = 4;
    RS_TRY(c.gemm_part.ensure(sizeof(double) * 3 * (size_t)nb * nl.nz));
    double *W = c.gemm_part.as<double>();
    hipLaunchKernelGGL((k_dproject<T, NBB>), dim3(nl.nz, (nb + NBB - 1) / NBB), dim3(256), 0,
                       c.stream, X, c.g.ML, nb, nl.nz, nl.d_ptr.as<int>(), nl.d_idx.as<int>(),
                       nl.d_val.as<double>(), occ_dev, W);
    RS_HIP(hipGetLastError());
    const int *aptr = nl.amap_dev.as<int>();
    const int *alma = aptr + (nl.natom + 1);
    hipLaunchKernelGGL((k_force_reduce<T>), dim3(nl.natom), dim3(256), 0, c.stream,
                       c.coefC.as<T>(), W, nb, nl.nz, aptr, alma, occ_dev, c.g.dV, force_out);
    RS_HIP(hipGetLastError());
    return kOK;
}

template int density_dev<double>(const double *, int, const double *, double *);
template int density_dev<float>(const float *, int, const double *, double *);
template int force_dev<double>(const double *, int, const double *, double *);
template int force_dev<float>(const float *, int, const double *, double *);

}  // namespace rsdft
