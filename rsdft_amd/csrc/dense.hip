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

#include <cmath>

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


// ---------------------------------------------------------------------------
// Eigen-decomposition of a nearly diagonal Hsub by iterative refinement.
//
// After the first sweeps the Ritz vectors change little from sweep to sweep,
// so Hsub (in the orthonormalised basis of the CG-updated Ritz vectors) is
// nearly diagonal.  Starting from X = I + E0, E0_ij = H_ij/(H_jj - H_ii), the
// refinement of Ogita and Aishima (Japan J. Indust. Appl. Math. 35, 2018)
//     R = I - X^T X,  S = X^T H X,  l_i = S_ii/(1 - R_ii),
//     E_ij = (S_ij + l_j R_ij)/(l_j - l_i)  (|l_i - l_j| > delta),  E_ii = R_ii/2,
//     X <- X (I + E),  delta = 2(|S - diag(l)|_F + |H|_F |R|_F)
// converges quadratically when the eigenvalues are separated by more than
// delta.  Each step is four n x n FP64 GEMMs instead of ~1,500 small kernels
// of the tridiagonal-reduction eigensolver.  Used only when every
// off-diagonal element is below 0.3 of its diagonal gap; if a pair stays
// inside delta or the iteration does not reach 1e-12 the eigensolve falls
// back to rocsolver_dsyevd.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// X = I + E0; st[0] = max off/gap ratio
__global__ void k_rs_init(const double *__restrict__ H, int n, double *__restrict__ X, double *st)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    double x = 0.0;
    if (i == j) {
        x = 1.0;
    } else {
        const double gap = H[(long)j * n + j] - H[(long)i * n + i];
        const double hij = H[(long)j * n + i];
        const double ratio = fabs(gap) > 0.0 ? fabs(hij) / fabs(gap) : (hij == 0.0 ? 0.0 : 1e300);
        atomic_max_nonneg(&st[0], ratio);
        x = gap != 0.0 ? hij / gap : 0.0;
    }
    X[(long)j * n + i] = x;
}

__global__ void k_set_identity(double *A, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) A[(long)j * n + i] = i == j ? 1.0 : 0.0;
}

// st[1] = |offdiag S|_F^2, st[2] = |R|_F^2, st[3] = |H|_F^2 (when hnorm)
__global__ void k_rs_norms(const double *__restrict__ S, const double *__restrict__ R,
                           const double *__restrict__ H, int n, int hnorm, double *st)
{
    __shared__ double red[3][16];
    double a = 0.0, b = 0.0, h = 0.0;
    const long nn = (long)n * n;
    for (long t = threadIdx.x; t < nn; t += blockDim.x) {
        const int j = (int)(t / n), i = (int)(t - (long)j * n);
        if (i != j) a += S[t] * S[t];
        b += R[t] * R[t];
        if (hnorm) h += H[t] * H[t];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        h += __shfl_down(h, off, 64);
    }
    if (lane == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = h; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) { s0 += red[0][k]; s1 += red[1][k]; s2 += red[2][k]; }
        st[1] = s0;
        st[2] = s1;
        if (hnorm) st[3] = s2;
    }
}

// F = I + E; lam_i = S_ii/(1 - R_ii); st[4] = number of pairs inside delta
__global__ void k_rs_update(const double *__restrict__ S, const double *__restrict__ R, int n,
                            double *__restrict__ F, double *__restrict__ lam, double *st)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double li = S[(long)i * n + i] / (1.0 - R[(long)i * n + i]);
    const double lj = S[(long)j * n + j] / (1.0 - R[(long)j * n + j]);
    if (i == j) lam[i] = li;
    const double delta = 2.0 * (sqrt(st[1]) + sqrt(st[3]) * sqrt(st[2]));
    double e;
    const double rij = R[(long)j * n + i];
    if (i == j) {
        e = 0.5 * rij;
    } else if (fabs(li - lj) > delta) {
        e = (S[(long)j * n + i] + lj * rij) / (lj - li);
    } else {
        e = 0.5 * rij;
        atomicAdd(&st[4], 1.0);
    }
    F[(long)j * n + i] = (i == j ? 1.0 : 0.0) + e;
}

// ascending order of lam (ties by index) -> V(:, rank) = X(:, i), D(rank) = lam(i)
__global__ void k_rs_sort(const double *__restrict__ X, const double *__restrict__ lam, int n,
                          double *__restrict__ V, double *__restrict__ D)
{
    const int i = blockIdx.x;
    const double li = lam[i];
    __shared__ int rk[256];
    int c = 0;
    for (int j = threadIdx.x; j < n; j += blockDim.x)
        c += (lam[j] < li || (lam[j] == li && j < i)) ? 1 : 0;
    rk[threadIdx.x] = c;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) rk[threadIdx.x] += rk[threadIdx.x + s];
        __syncthreads();
    }
    const int rank = rk[0];
    if (threadIdx.x == 0) D[rank] = li;
    for (int k = threadIdx.x; k < n; k += blockDim.x) V[(long)rank * n + k] = X[(long)i * n + k];
}

static int read_st(double *host, const double *dev, int n)
{
    Context &c = ctx();
    RS_HIP(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

// On success V (n x n) and D (n) hold the eigenpairs of H, *ok = 1.
static int refine_eigh(const double *H, int n, double *V, double *D, int *ok)
{
    Context &c = ctx();
    *ok = 0;
    const size_t nn = (size_t)n * n;
    RS_TRY(c.evwork.ensure(sizeof(double) * (6 * nn + n + 8)));
    double *X = c.evwork.as<double>(), *B = X + nn, *S = B + nn, *R = S + nn, *F = R + nn,
           *Xn = F + nn, *lam = Xn + nn, *st = lam + n;
    const dim3 g2((n + 127) / 128, n), b2(128);
    RS_HIP(hipMemsetAsync(st, 0, sizeof(double) * 8, c.stream));
    hipLaunchKernelGGL(k_rs_init, g2, b2, 0, c.stream, H, n, X, st);
    double hs[8];
    RS_TRY(read_st(hs, st, 8));
    if (!(hs[0] < 0.3)) return kOK;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    bool converged = false;
    for (int it = 0; it < 10; ++it) {
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                              H, n, X, n, &zero, B, n));
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, n,
                              &one, X, n, B, n, &zero, S, n));
        hipLaunchKernelGGL(k_set_identity, g2, b2, 0, c.stream, R, n);
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, n,
                              &mone, X, n, X, n, &one, R, n));
        RS_HIP(hipMemsetAsync(st + 4, 0, sizeof(double), c.stream));
        hipLaunchKernelGGL(k_rs_norms, dim3(1), dim3(1024), 0, c.stream, S, R, H, n, it == 0 ? 1 : 0, st);
        hipLaunchKernelGGL(k_rs_update, g2, b2, 0, c.stream, S, R, n, F, lam, st);
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one,
                              X, n, F, n, &zero, Xn, n));
        std::swap(X, Xn);
        RS_TRY(read_st(hs, st, 8));
        const double hn = sqrt(hs[3]);
        if (hs[4] > 0.0 && it >= 3) break;                       // unresolved cluster
        if (sqrt(hs[1]) <= 1e-12 * hn && sqrt(hs[2]) <= 1e-12) { converged = hs[4] == 0.0; break; }
    }
    if (!converged) return kOK;
    hipLaunchKernelGGL(k_rs_sort, dim3(n), dim3(256), 0, c.stream, X, lam, n, V, D);
    RS_HIP(hipGetLastError());
    *ok = 1;
    return kOK;
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
    int refined = 0;
    double *Vec = H;
    if (c.eig_mode == 1) {
        RS_TRY(c.mat3.ensure(sizeof(double) * (size_t)nb * nb));
        RS_TRY(refine_eigh(H, nb, c.mat3.as<double>(), D, &refined));
        if (refined) Vec = c.mat3.as<double>();
        c.eig_refined += refined;
    }
    if (!refined) {
        RS_BLAS(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, nb, H, nb, D,
                                 D + nb, info));
        hipLaunchKernelGGL(k_check_info, dim3(1), dim3(1), 0, c.stream, info, info + 3);
        hipLaunchKernelGGL(k_fix_signs, dim3(nb), dim3(128), 0, c.stream, H, nb);
    }
    T *Out = c.tmp.as<T>();
    RS_TRY(gemm_nn<T>(X, Vec, nb, Out));
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
