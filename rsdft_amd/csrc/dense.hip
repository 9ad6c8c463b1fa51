// Dense band-space algebra: Gram-Schmidt and subspace diagonalisation.
//
// gram_schmidt_t (gram_schmidt_t_module.f90:84-316) is single-pass blocked
// classical Gram-Schmidt with positive normalisation; in exact arithmetic it
// returns Q = Psi R^-1 with R^T R = dV Psi^T Psi, R upper triangular with a
// positive diagonal.  Here that is computed as Cholesky-QR: one tall-skinny
// GEMM for S, an FP64 Cholesky and triangular inverse on the GPU, one GEMM
// for the update.  subspace_diag_la (subspace_diag_la_module.f90:85-206)
// becomes H Psi (batched), one GEMM for Psi^T H Psi, the dsyr2k
// symmetrisation, an FP64 eigensolve on the GPU and one rotation GEMM.
// Plain GEMMs go to rocBLAS (atomics disabled, so results are
// deterministic); the FP32-storage GEMMs are hand-written MFMA kernels
// (FP32 or BF16 hi+lo products, FP64-flushed accumulation).
#include "common.h"

namespace rsdft {

template <class T>
int split_gemm_tn(const float *A, const float *B, int nb, long K, double alpha, double *C);
template <class T>
int split_gemm_nn(const float *A, const double *U, int nb, long M, float *C);

template <class S, class D>
__global__ void k_convert(const S *src, D *dst, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = (D)src[i];
}

template <class T>
int convert_from_f64(const double *src, T *dst, long n)
{
    Context &c = ctx();
    if ((const void *)src == (const void *)dst) return kOK;
    hipLaunchKernelGGL((k_convert<double, T>), dim3(2048), dim3(256), 0, c.stream, src, dst, n);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int convert_to_f64(const T *src, double *dst, long n)
{
    Context &c = ctx();
    if ((const void *)src == (const void *)dst) return kOK;
    hipLaunchKernelGGL((k_convert<T, double>), dim3(2048), dim3(256), 0, c.stream, src, dst, n);
    RS_HIP(hipGetLastError());
    return kOK;
}

// Hsub = zz (C + C^T), zz = 0.5 dV (dsyr2k('U','T',...), subspace_diag_la_module.f90:96)
__global__ void k_symmetrize(const double *C, double *H, int n, double zz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= n) return;
    H[i + (long)j * n] = zz * (C[i + (long)j * n] + C[j + (long)i * n]);
}

// zero the strictly lower triangle (column-major)
__global__ void k_zero_lower(double *A, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i < n && i > j) A[i + (long)j * n] = 0.0;
}

// eigenvector sign convention: V(j,j) >= 0 (any sign is an eigenvector; this
// one keeps V close to I for a converging sweep)
__global__ void k_fix_signs(double *V, int n)
{
    const int j = blockIdx.x;
    const bool flip = V[(long)j * n + j] < 0.0;
    if (!flip) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) V[(long)j * n + i] = -V[(long)j * n + i];
}

__global__ void k_check_info(const int *info, int *flag)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && *info != 0) *flag = *info;
}

// C(nb x nb) = alpha * A^T B, A and B (ML x nb) column-major, result FP64
template <class T>
static int gemm_tn(const T *A, const T *B, int nb, double alpha, double *C)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    if (sizeof(T) == 8) {
        const double a = alpha, z = 0.0;
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb,
                              (int)ML, &a, (const double *)A, (int)ML, (const double *)B, (int)ML,
                              &z, C, nb));
        return kOK;
    }
    return split_gemm_tn<T>((const float *)A, (const float *)B, nb, ML, alpha, C);
}

// Out(ML x nb) = A(ML x nb) * U(nb x nb), U FP64
template <class T>
static int gemm_nn(const T *A, const double *U, int nb, T *Out)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    if (sizeof(T) == 8) {
        const double a = 1.0, z = 0.0;
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, (int)ML, nb,
                              nb, &a, (const double *)A, (int)ML, U, nb, &z, (double *)Out,
                              (int)ML));
        return kOK;
    }
    return split_gemm_nn<T>((const float *)A, U, nb, ML, (float *)Out);
}

static int ensure_dense(int nb)
{
    Context &c = ctx();
    RS_TRY(c.mat.ensure(sizeof(double) * (size_t)nb * nb));
    RS_TRY(c.mat2.ensure(sizeof(double) * (size_t)nb * nb));
    RS_TRY(c.eval.ensure(sizeof(double) * (size_t)nb * 2));
    RS_TRY(c.einfo.ensure(sizeof(int) * 4));
    return kOK;
}

// psi <- psi R^-1 with R^T R = dV psi^T psi, into c.tmp, then swap with c.psi
template <class T>
int gram_schmidt_dev(T *X, int nb)
{
    Context &c = ctx();
    RS_TRY(ensure_dense(nb));
    double *S = c.mat.as<double>();
    int *info = c.einfo.as<int>();
    RS_TRY(gemm_tn<T>(X, X, nb, c.g.dV, S));
    RS_BLAS(rocsolver_dpotrf(c.blas, rocblas_fill_upper, nb, S, nb, info));
    hipLaunchKernelGGL(k_check_info, dim3(1), dim3(1), 0, c.stream, info, info + 2);
    RS_BLAS(rocsolver_dtrtri(c.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, nb, S, nb,
                             info + 1));
    hipLaunchKernelGGL(k_zero_lower, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, S, nb);
    RS_HIP(hipGetLastError());
    T *Out = c.tmp.as<T>();
    RS_TRY(gemm_nn<T>(X, S, nb, Out));
    std::swap(c.psi.p, c.tmp.p);
    std::swap(c.psi.bytes, c.tmp.bytes);
    return kOK;
}

// Rayleigh-Ritz in place (psi rotated into c.tmp then swapped), esp ascending
template <class T>
int subspace_diag_dev(T *X, int nb, double *esp_dev)
{
    Context &c = ctx();
    RS_TRY(ensure_dense(nb));
    T *HX = c.hp.as<T>();
    RS_TRY(hpsi_apply<T>(X, HX, nb, kKin | kLoc | kNl));
    double *C = c.mat2.as<double>(), *H = c.mat.as<double>();
    RS_TRY(gemm_tn<T>(X, HX, nb, 1.0, C));
    hipLaunchKernelGGL(k_symmetrize, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, C, H, nb,
                       0.5 * c.g.dV);
    RS_HIP(hipGetLastError());
    int *info = c.einfo.as<int>();
    double *D = c.eval.as<double>();
    RS_BLAS(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, nb, H, nb, D,
                             D + nb, info));
    hipLaunchKernelGGL(k_check_info, dim3(1), dim3(1), 0, c.stream, info, info + 3);
    hipLaunchKernelGGL(k_fix_signs, dim3(nb), dim3(128), 0, c.stream, H, nb);
    T *Out = c.tmp.as<T>();
    RS_TRY(gemm_nn<T>(X, H, nb, Out));
    RS_HIP(hipMemcpyAsync(esp_dev, D, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
    std::swap(c.psi.p, c.tmp.p);
    std::swap(c.psi.bytes, c.tmp.bytes);
    return kOK;
}

template int convert_from_f64<double>(const double *, double *, long);
template int convert_from_f64<float>(const double *, float *, long);
template int convert_to_f64<double>(const double *, double *, long);
template int convert_to_f64<float>(const float *, double *, long);
template int gram_schmidt_dev<double>(double *, int);
template int gram_schmidt_dev<float>(float *, int);
template int subspace_diag_dev<double>(double *, int, double *);
template int subspace_diag_dev<float>(float *, int, double *);

}  // namespace rsdft
