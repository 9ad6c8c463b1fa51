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
// The tall A^T B products are hand-written split-K MFMA kernels (FP64, or
// FP32 / BF16 hi+lo products with FP64-flushed accumulation); the FP64
// rotations and the small n x n products go to rocBLAS (atomics disabled, so
// results are deterministic).
#include "common.h"

#include <cmath>

namespace rsdft {

int eigh_tridiag(const double *H, int n, double *V, double *D, int *einfo, int *done,
                 int tri_solver);
template <class T>
int split_gemm_tn(const float *A, const float *B, int nb, long K, double alpha, double *C,
                  int sym_result);
template <class T>
int split_gemm_nn(const float *A, const double *U, int nb, long M, float *C);
int mfma_gemm_tn_f64(const double *A, const double *B, int nb, long K, double alpha, double *C,
                     int sym_result);

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
// nearly diagonal, except for groups of nearly degenerate bands that CG mixes
// strongly.  Starting from X = I, each iteration
//   1. finds the clusters: connected components of the pairs whose
//      first-order correction |E_ij| = |B_ij + l_j R_ij| / |l_j - l_i|
//      exceeds 0.1 (B = X^T H X, R = I - X^T X, l_i = B_ii / (1 - R_ii)),
//   2. diagonalises every cluster block of B exactly (parallel cyclic Jacobi
//      in LDS, one workgroup per cluster) and rotates X by the block-diagonal
//      result Q,
//   3. takes one refinement step of Ogita and Aishima (Japan J. Indust. Appl.
//      Math. 35, 2018) for all remaining pairs:
//         E_ij = (B_ij + l_j R_ij)/(l_j - l_i) between clusters,
//         E_ij = R_ij / 2 inside a cluster and on the diagonal,
//         X <- X (I + E),
// and stops when max |B_ij| (i != j) <= 1e-12 max |H| and max |R_ij| <= 1e-12, or
// when both are below 1e-10 and stopped shrinking (the floor eps / gap of a
// close pair outside every cluster); the last X is then orthonormalised by
// X <- X (I + R/2), which leaves |X^T X - I| ~ |R|^2.
// Every step is an n x n FP64 GEMM or a single small kernel, instead of the
// ~1,500 small kernels of the tridiagonal-reduction eigensolver.  A cluster
// larger than kClusterStart at the start (kClusterMax later), too many strong
// pairs, or no quadratic convergence within kRefineIters iterations falls back
// to rocsolver_dsyevd.
// ---------------------------------------------------------------------------
constexpr int kClusterMax = 48;
constexpr int kPairMax = 2048;
constexpr int kRefineIters = 8;
constexpr double kStrongPair = 0.1;
constexpr int kClusterStart = 16;   // largest cluster accepted at the first iteration

__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

__global__ void k_set_identity(double *A, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) A[(long)j * n + i] = i == j ? 1.0 : 0.0;
}

// F = I + R / 2
__global__ void k_rs_half_step(const double *__restrict__ R, int n, double *__restrict__ F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) F[(long)j * n + i] = (i == j ? 1.0 : 0.0) + 0.5 * R[(long)j * n + i];
}

// st[1] = max |offdiag B|, st[2] = max |R|, st[3] = max |H| (when hnorm); R may
// be null.  Grid-stride over the n^2 entries; every workgroup folds its maxima
// into st by an atomic max on the bit patterns (non-negative doubles order like
// them), so those entries must be zero on entry (k_rs_clear_norms).
__global__ void __launch_bounds__(256)
k_rs_norms(const double *__restrict__ S, const double *__restrict__ R,
           const double *__restrict__ H, int n, int hnorm, double *st)
{
    __shared__ double red[3][4];
    double a = 0.0, b = 0.0, h = 0.0;
    const int nn = n * n;   // n <= 46340
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += gridDim.x * blockDim.x) {
        if (t % (n + 1) != 0) a = fmax(a, fabs(S[t]));   // off the diagonal
        if (R) b = fmax(b, fabs(R[t]));
        if (hnorm) h = fmax(h, fabs(H[t]));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_down(a, off, 64));
        b = fmax(b, __shfl_down(b, off, 64));
        h = fmax(h, __shfl_down(h, off, 64));
    }
    if (lane == 0) { red[0][w] = a; red[1][w] = b; red[2][w] = h; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
            s0 = fmax(s0, red[0][k]);
            s1 = fmax(s1, red[1][k]);
            s2 = fmax(s2, red[2][k]);
        }
        atomic_max_nonneg(st + 1, s0);
        atomic_max_nonneg(st + 2, s1);
        if (hnorm) atomic_max_nonneg(st + 3, s2);
    }
}

__global__ void k_rs_clear_norms(double *st, int hnorm)
{
    if (threadIdx.x == 0) {
        st[1] = 0.0;
        st[2] = 0.0;
        if (hnorm) st[3] = 0.0;
    }
}

// k_rs_norms over enough workgroups to spread the n^2 entries
static void rs_norms(const double *S, const double *R, const double *H, int n, int hnorm,
                     double *st)
{
    Context &c = ctx();
    const int nb = (int)std::min<long>(128, ((long)n * n + 255) / 256);
    hipLaunchKernelGGL(k_rs_clear_norms, dim3(1), dim3(64), 0, c.stream, st, hnorm);
    hipLaunchKernelGGL(k_rs_norms, dim3(nb), dim3(256), 0, c.stream, S, R, H, n, hnorm, st);
}

// Strong pairs i < j of the cluster search (|B_ij + l_j R_ij| > kStrongPair
// |l_j - l_i|, l_i = B_ii / (1 - R_ii)) over many workgroups, appended to
// pairs[0 .. kPairMax) (i) and pairs[kPairMax ..) (j) through the counter
// pairs[2 kPairMax], which must be zero on entry; the count may exceed
// kPairMax (then only the first kPairMax are stored).
__global__ void __launch_bounds__(256)
k_rs_pairs(const double *__restrict__ B, const double *__restrict__ R, int n,
           int *__restrict__ pairs)
{
    const int nn = n * n;   // n <= 46340
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += gridDim.x * blockDim.x) {
        const int j = t / n, i = t - j * n;
        if (i >= j) continue;
        const double li = B[(long)i * n + i] / (1.0 - (R ? R[(long)i * n + i] : 0.0));
        const double lj = B[(long)j * n + j] / (1.0 - (R ? R[(long)j * n + j] : 0.0));
        const double num = B[t] + lj * (R ? R[t] : 0.0);
        if (fabs(num) > kStrongPair * fabs(lj - li)) {
            const int k = atomicAdd(pairs + 2 * kPairMax, 1);
            if (k < kPairMax) {
                pairs[k] = i;
                pairs[kPairMax + k] = j;
            }
        }
    }
}

__global__ void k_rs_clear_pairs(int *pairs)
{
    if (threadIdx.x == 0) pairs[2 * kPairMax] = 0;
}

// Cluster search, one workgroup.  lam_i = B_ii/(1 - R_ii); strong pairs;
// label propagation to connected components; compact member lists:
// cl[0] = number of clusters (size >= 2), cl[1 .. ncl+1] = offsets,
// cl[n+2 ..] = members.  clu[i] = cluster root of i (i itself if single).
// st[5] = 1 on overflow (pairs or cluster size), st[6] = max cluster size,
// st[7] = number of strong pairs.  R may be null (R = 0).
__global__ void __launch_bounds__(1024)
k_rs_clusters(const double *__restrict__ B, const double *__restrict__ R, int n,
              const int *__restrict__ pairs, double *__restrict__ lam, int *__restrict__ clu,
              int *__restrict__ cl, double *st)
{
    extern __shared__ unsigned char smem_raw[];
    double *l = reinterpret_cast<double *>(smem_raw);
    int *label = reinterpret_cast<int *>(l + n);
    int *cnt = label + n;
    int *pi = cnt + n;
    int *pj = pi + kPairMax;
    __shared__ int npairs, changed, ncl, fail, maxsz;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < n; i += nt) {
        const double rii = R ? R[(long)i * n + i] : 0.0;
        l[i] = B[(long)i * n + i] / (1.0 - rii);
        lam[i] = l[i];
        label[i] = i;
        cnt[i] = 0;
    }
    if (tid == 0) { npairs = pairs[2 * kPairMax]; fail = 0; maxsz = 0; }
    __syncthreads();
    for (int k = tid; k < min(npairs, kPairMax); k += nt) {
        pi[k] = pairs[k];
        pj[k] = pairs[kPairMax + k];
    }
    __syncthreads();
    const int np = npairs < kPairMax ? npairs : kPairMax;
    if (tid == 0 && npairs > kPairMax) fail = 1;
    do {
        __syncthreads();
        if (tid == 0) changed = 0;
        __syncthreads();
        for (int k = tid; k < np; k += nt) {
            const int a = label[pi[k]], b = label[pj[k]];
            if (a != b) {
                const int m = a < b ? a : b;
                atomicMin(&label[pi[k]], m);
                atomicMin(&label[pj[k]], m);
                atomicMin(&label[a], m);
                atomicMin(&label[b], m);
                changed = 1;
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += nt) {   // pointer jumping
            const int a = label[i], b = label[a];
            if (b < a) { label[i] = b; changed = 1; }
        }
        __syncthreads();
    } while (changed);
    for (int i = tid; i < n; i += nt) atomicAdd(&cnt[label[i]], 1);
    __syncthreads();
    if (tid == 0) {
        // roots of clusters of size >= 2 in index order; cnt[root] becomes its offset
        int off = 0, c = 0;
        for (int i = 0; i < n; ++i) {
            const int sz = cnt[i];
            if (label[i] == i && sz >= 2) {
                if (sz > maxsz) maxsz = sz;
                if (sz > kClusterMax) fail = 1;
                cl[1 + c] = off;
                cnt[i] = off;
                off += sz;
                ++c;
            } else {
                cnt[i] = -1;
            }
        }
        cl[1 + c] = off;
        cl[0] = c;
        ncl = c;
        st[5] = fail ? 1.0 : 0.0;
        st[6] = maxsz;
        st[7] = npairs;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        const int r = label[i];
        clu[i] = r;
        if (cnt[r] >= 0) {
            const int pos = atomicAdd(&cnt[r], 1);
            cl[n + 2 + pos] = i;
        }
    }
}

// Members of one cluster in ascending index order (atomics scatter them).
__device__ void sort_members(int *idx, int m)
{
    if (threadIdx.x == 0)
        for (int a = 1; a < m; ++a) {
            const int v = idx[a];
            int b = a - 1;
            while (b >= 0 && idx[b] > v) { idx[b + 1] = idx[b]; --b; }
            idx[b + 1] = v;
        }
    __syncthreads();
}

// Q(cluster, cluster) = eigenvectors of B(cluster, cluster) by parallel cyclic
// Jacobi (round-robin pairing, m/2 disjoint rotations per round).  Q must be
// the identity outside the clusters.  st[4] += sweeps that did not converge.
__global__ void __launch_bounds__(256)
k_rs_jacobi(const double *__restrict__ B, int n, const int *__restrict__ cl, double *__restrict__ Q,
            double *st)
{
    __shared__ double A[kClusterMax][kClusterMax + 1];
    __shared__ double V[kClusterMax][kClusterMax + 1];
    __shared__ double cs[kClusterMax / 2], sn[kClusterMax / 2];
    __shared__ int pp[kClusterMax / 2], qq[kClusterMax / 2];
    __shared__ int idx[kClusterMax];
    __shared__ double offn, diagn;
    const int ncl = cl[0];
    if ((int)blockIdx.x >= ncl) return;
    const int b0 = cl[1 + blockIdx.x], m = cl[2 + blockIdx.x] - b0;
    if (m > kClusterMax) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int a = tid; a < m; a += nt) idx[a] = cl[n + 2 + b0 + a];
    __syncthreads();
    sort_members(idx, m);
    for (int t = tid; t < m * m; t += nt) {
        const int a = t % m, b = t / m;
        A[a][b] = B[(long)idx[b] * n + idx[a]];
        V[a][b] = a == b ? 1.0 : 0.0;
    }
    const int me = m + (m & 1);   // even size; index m (odd m) is a dummy
    const int np = me / 2;
    int sweep = 0;
    for (; sweep < 40; ++sweep) {
        __syncthreads();
        if (tid == 0) { offn = 0.0; diagn = 0.0; }
        __syncthreads();
        double o = 0.0, d = 0.0;
        for (int t = tid; t < m * m; t += nt) {
            const int a = t % m, b = t / m;
            if (a != b) o += A[a][b] * A[a][b];
            else d += A[a][b] * A[a][b];
        }
        atomicAdd(&offn, o);
        atomicAdd(&diagn, d);
        __syncthreads();
        if (offn <= 1e-30 * diagn || offn == 0.0) break;
        for (int rnd = 0; rnd < me - 1; ++rnd) {
            if (tid < np) {
                // round-robin: position 0 fixed, the others rotate
                const int k = tid;
                int p = k == 0 ? 0 : 1 + (k - 1 + rnd) % (me - 1);
                int q = 1 + (me - 2 - k + rnd) % (me - 1);
                if (p > q) { const int t2 = p; p = q; q = t2; }
                double c = 1.0, s = 0.0;
                if (q < m && A[p][q] != 0.0) {
                    const double apq = A[p][q];
                    const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    s = t * c;
                }
                pp[k] = p; qq[k] = q; cs[k] = c; sn[k] = s;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J
            for (int t = tid; t < np * m; t += nt) {
                const int k = t / m, r = t - k * m;
                const int p = pp[k], q = qq[k];
                if (q >= m) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[r][p], aq = A[r][q];
                A[r][p] = c * ap - s * aq;
                A[r][q] = s * ap + c * aq;
                const double vp = V[r][p], vq = V[r][q];
                V[r][p] = c * vp - s * vq;
                V[r][q] = s * vp + c * vq;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int t = tid; t < np * m; t += nt) {
                const int k = t / m, r = t - k * m;
                const int p = pp[k], q = qq[k];
                if (q >= m) continue;
                const double c = cs[k], s = sn[k];
                const double ap = A[p][r], aq = A[q][r];
                A[p][r] = c * ap - s * aq;
                A[q][r] = s * ap + c * aq;
            }
            __syncthreads();
        }
    }
    if (tid == 0 && sweep >= 40) atomicAdd(&st[4], 1.0);
    for (int t = tid; t < m * m; t += nt) {
        const int a = t % m, b = t / m;
        Q[(long)idx[b] * n + idx[a]] = V[a][b];
    }
}

// F = I + E of the refinement step; lam_i = B_ii/(1 - R_ii)
__global__ void k_rs_update(const double *__restrict__ S, const double *__restrict__ R, int n,
                            const int *__restrict__ clu, double *__restrict__ F,
                            double *__restrict__ lam)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double li = S[(long)i * n + i] / (1.0 - R[(long)i * n + i]);
    const double lj = S[(long)j * n + j] / (1.0 - R[(long)j * n + j]);
    if (i == j) lam[i] = li;
    const double rij = R[(long)j * n + i];
    double e;
    if (i == j || clu[i] == clu[j]) {
        e = 0.5 * rij;
    } else {
        e = (S[(long)j * n + i] + lj * rij) / (lj - li);
        // a pair that became strong after the cluster rotation waits for the
        // next cluster search instead of taking a large first-order step
        if (!(fabs(e) <= 0.3)) e = 0.5 * rij;
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

// B = X^T H X, R = I - X^T X (T is scratch)
static int rs_project(const double *H, const double *X, int n, double *T, double *B, double *R)
{
    Context &c = ctx();
    const double one = 1.0, zero = 0.0, mone = -1.0;
    const dim3 g2((n + 127) / 128, n), b2(128);
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, H,
                          n, X, n, &zero, T, n));
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, n, &one,
                          X, n, T, n, &zero, B, n));
    hipLaunchKernelGGL(k_set_identity, g2, b2, 0, c.stream, R, n);
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, n,
                          &mone, X, n, X, n, &one, R, n));
    return kOK;
}

// On success V (n x n) and D (n) hold the eigenpairs of H, *ok = 1.
static int refine_eigh(const double *H, int n, double *V, double *D, int *ok)
{
    Context &c = ctx();
    *ok = 0;
    if (n < 2) return kOK;   // nothing to refine (and no pair for the Jacobi launch)
    const size_t nn = (size_t)n * n;
    const size_t lds = (size_t)n * sizeof(double) + 2 * (size_t)n * sizeof(int) +
                       2 * (size_t)kPairMax * sizeof(int);
    if (lds > 64 * 1024) return kOK;   // cluster search does not fit one workgroup
    RS_TRY(c.evwork.ensure(sizeof(double) * (6 * nn + n + 8) +
                           sizeof(int) * (3 * n + 4 + 2 * kPairMax + 1)));
    double *X = c.evwork.as<double>(), *T = X + nn, *B = T + nn, *R = B + nn, *F = R + nn,
           *Xn = F + nn, *lam = Xn + nn, *st = lam + n;
    int *clu = reinterpret_cast<int *>(st + 8), *cl = clu + n;
    int *pairs = cl + 2 * n + 4;   // 2 kPairMax + 1 ints
    const dim3 g2((n + 127) / 128, n), b2(128);
    const double one = 1.0, zero = 0.0;
    RS_HIP(hipMemsetAsync(st, 0, sizeof(double) * 8, c.stream));
    rs_norms(H, nullptr, H, n, 1, st);
    hipLaunchKernelGGL(k_set_identity, g2, b2, 0, c.stream, X, n);
    const double *Bc = H, *Rc = nullptr;   // B and R of the current X (X = I)
    double hs[8];
    bool converged = false;
    double prev_off = -1.0;
    for (int it = 0; it < kRefineIters; ++it) {
        // 1-2: exact rotation inside every cluster
        hipLaunchKernelGGL(k_rs_clear_pairs, dim3(1), dim3(64), 0, c.stream, pairs);
        hipLaunchKernelGGL(k_rs_pairs, dim3((unsigned)std::min<long>(256, ((long)n * n + 255) / 256)),
                           dim3(256), 0, c.stream, Bc, Rc, n, pairs);
        hipLaunchKernelGGL(k_rs_clusters, dim3(1), dim3(1024), lds, c.stream, Bc, Rc, n, pairs,
                           lam, clu, cl, st);
        if (it == 0) {
            // large clusters at the start mean the Ritz vectors are still moving:
            // the tridiagonal eigensolver is cheaper than many refinement steps
            RS_TRY(read_st(hs, st, 8));
            if (hs[5] != 0.0 || hs[6] > kClusterStart) return kOK;
        }
        hipLaunchKernelGGL(k_set_identity, g2, b2, 0, c.stream, F, n);
        hipLaunchKernelGGL(k_rs_jacobi, dim3(n / 2), dim3(256), 0, c.stream, Bc, n, cl, F, st);
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n,
                              &one, X, n, F, n, &zero, Xn, n));
        std::swap(X, Xn);
        RS_TRY(rs_project(H, X, n, T, B, R));
        // 3: refinement step between clusters
        hipLaunchKernelGGL(k_rs_update, g2, b2, 0, c.stream, B, R, n, clu, F, lam);
        RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n,
                              &one, X, n, F, n, &zero, Xn, n));
        std::swap(X, Xn);
        RS_TRY(rs_project(H, X, n, T, B, R));
        Bc = B;
        Rc = R;
        rs_norms(B, R, H, n, 0, st);
        RS_HIP(hipGetLastError());
        RS_TRY(read_st(hs, st, 8));
        if (hs[5] != 0.0 || hs[4] != 0.0) return kOK;   // cluster too large / Jacobi stalled
        // rounding level: the next Gram-Schmidt restores orthogonality below 1e-12 anyway
        if (hs[1] <= 1e-12 * hs[3] && hs[2] <= 1e-12) { converged = true; break; }
        // a pair of gap d keeps |R| near eps / d (each step's correction divides
        // rounding by d): below 1e-10 and no longer shrinking, the iteration
        // has reached that floor
        if (prev_off >= 0.0 && hs[1] > 0.25 * prev_off && hs[1] <= 1e-10 * hs[3] &&
            hs[2] <= 1e-10) { converged = true; break; }
        if (it >= 2 && hs[1] > 1e-5 * hs[3]) return kOK;   // not converging quadratically
        prev_off = hs[1];
    }
    if (!converged) return kOK;
    // eigenvalues: diagonal of the final B
    hipLaunchKernelGGL(k_rs_update, g2, b2, 0, c.stream, B, R, n, clu, F, lam);
    // symmetric (Loewdin) orthonormalisation to first order, X <- X (I + R/2)
    // with R = I - X^T X: the orthogonality defect goes from |R| to |R|^2
    hipLaunchKernelGGL(k_rs_half_step, g2, b2, 0, c.stream, R, n, F);
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, n, n, n,
                          &one, X, n, F, n, &zero, Xn, n));
    std::swap(X, Xn);
    hipLaunchKernelGGL(k_rs_sort, dim3(n), dim3(256), 0, c.stream, X, lam, n, V, D);
    RS_HIP(hipGetLastError());
    *ok = 1;
    return kOK;
}

// C(nb x nb) = alpha * A^T B, A and B (ML x nb) column-major, result FP64,
// on the hand-written split-K MFMA kernels in both precisions (rocBLAS dgemm
// computes the full product; dsyrk for A == B measured ~10x slower than dgemm
// at 480 x 373,248).  sym_result (or A == B): A^T B is symmetric in exact
// arithmetic; only the tile pairs i <= j are computed and mirrored
template <class T>
static int gemm_tn(const T *A, const T *B, int nb, double alpha, double *C,
                   bool sym_result = false)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    if (sizeof(T) == 8)
        return mfma_gemm_tn_f64((const double *)A, (const double *)B, nb, ML, alpha, C,
                                sym_result ? 1 : 0);
    return split_gemm_tn<T>((const float *)A, (const float *)B, nb, ML, alpha, C,
                            sym_result ? 1 : 0);
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
    RS_TRY(c.mat3.ensure(sizeof(double) * (size_t)nb * nb));
    RS_TRY(c.mat4.ensure(sizeof(double) * (size_t)nb * nb));
    RS_TRY(c.eval.ensure(sizeof(double) * (size_t)nb * 2));
    RS_TRY(c.einfo.ensure(sizeof(int) * 4));
    return kOK;
}

// S <- R^-1 in place, S = dV X^T X = R^T R on entry (upper Cholesky factor)
// ---------------------------------------------------------------------------
// Upper Cholesky S = R^T R (dpotrf('U')) blocked by kPotfBlk columns: each
// diagonal block is factored by one workgroup with the block in LDS (right-
// looking, the reciprocal of the pivot scaling row j as dpotf2 does), then
// the block row to its right is solved (dtrsm, R12 = R11^-T S12) and the
// trailing matrix updated (dsyrk, S22 -= R12^T R12) by rocBLAS.  *info = j+1
// for the first non-positive pivot j (as dpotrf), and the factorisation stops.
// ---------------------------------------------------------------------------
constexpr int kPotfBlk = 120;
constexpr int kPotfThreads = 1024;

__global__ void k_clear_int(int *p) { *p = 0; }

__global__ void __launch_bounds__(kPotfThreads)
k_potf2_lds(double *__restrict__ A, int lda, int kb, int k0, int *__restrict__ info)
{
    extern __shared__ double Lb[];   // Lb[c * (kb + 1) + r] = A(r, c)
    __shared__ int bad;
    const int P = kb + 1, tid = threadIdx.x, nt = blockDim.x;
    constexpr int TX = 32, TY = kPotfThreads / 32;
    if (*info != 0) return;          // an earlier block failed (uniform)
    for (int e = tid; e < kb * kb; e += nt) {
        const int col = e / kb, r = e - col * kb;
        Lb[col * P + r] = A[(long)col * lda + r];
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    int jend = kb;
    for (int j = 0; j < kb; ++j) {
        const double d = Lb[j * P + j];
        if (!(d > 0.0)) {            // the same value in every thread
            if (tid == 0) bad = k0 + j + 1;
            jend = j;
            break;
        }
        const double rjj = sqrt(d), inv = 1.0 / rjj;
        __syncthreads();             // every thread has read the pivot
        for (int k = j + tid; k < kb; k += nt) Lb[k * P + j] = k == j ? rjj : Lb[k * P + j] * inv;
        __syncthreads();
        // trailing update on a 32 x TY thread grid (no index division)
        for (int k = j + 1 + tid / TX; k < kb; k += TY) {
            const double rk = Lb[k * P + j];
            for (int i = j + 1 + tid % TX; i <= k; i += TX) Lb[k * P + i] -= Lb[i * P + j] * rk;
        }
        __syncthreads();
    }
    for (int e = tid; e < kb * kb; e += nt) {
        const int col = e / kb, r = e - col * kb;
        if (r <= col && col < jend) A[(long)col * lda + r] = Lb[col * P + r];
    }
    if (tid == 0 && bad) *info = bad;
}

static int potrf_upper(double *S, int n, int *info)
{
    Context &c = ctx();
    static bool lds_set = false;
    const size_t lds = sizeof(double) * (size_t)kPotfBlk * (kPotfBlk + 1);
    if (!lds_set) {
        RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_potf2_lds),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        lds_set = true;
    }
    hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(1), 0, c.stream, info);
    const double one = 1.0, mone = -1.0;
    for (int k0 = 0; k0 < n; k0 += kPotfBlk) {
        const int kb = std::min(kPotfBlk, n - k0), rest = n - k0 - kb;
        double *A11 = S + (long)k0 * n + k0;
        hipLaunchKernelGGL(k_potf2_lds, dim3(1), dim3(kPotfThreads), lds, c.stream, A11, n, kb, k0,
                           info);
        RS_HIP(hipGetLastError());
        if (rest > 0) {
            double *A12 = S + (long)(k0 + kb) * n + k0;
            double *A22 = S + (long)(k0 + kb) * n + k0 + kb;
            RS_BLAS(rocblas_dtrsm(c.blas, rocblas_side_left, rocblas_fill_upper,
                                  rocblas_operation_transpose, rocblas_diagonal_non_unit, kb, rest,
                                  &one, A11, n, A12, n));
            RS_BLAS(rocblas_dsyrk(c.blas, rocblas_fill_upper, rocblas_operation_transpose, rest, kb,
                                  &mone, A12, n, &one, A22, n));
        }
    }
    return kOK;
}

static int chol_inverse(double *S, int nb)
{
    Context &c = ctx();
    int *info = c.einfo.as<int>();
    RS_TRY(potrf_upper(S, nb, info));
    hipLaunchKernelGGL(k_check_info, dim3(1), dim3(1), 0, c.stream, info, info + 2);
    RS_BLAS(rocsolver_dtrtri(c.blas, rocblas_fill_upper, rocblas_diagonal_non_unit, nb, S, nb,
                             info + 1));
    hipLaunchKernelGGL(k_zero_lower, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, S, nb);
    RS_HIP(hipGetLastError());
    return kOK;
}

// Eigenpairs of the symmetric Hsub H (destroyed): eigenvalues ascending in D,
// eigenvectors (columns, V(j,j) >= 0) at *Vec, which is H or Vbuf.
// eig_mode 1: refinement when H is nearly diagonal, else as mode 3;
// 2: the LDS-resident tridiagonal reduction with rocsolver_dstedc;
// 3: the reduction with bisection + inverse iteration on the tridiagonal;
// 0 (and matrices too large for the reduction): rocsolver_dsyevd.
int eigh_sub(double *H, int nb, double *Vbuf, double *D, double **Vec)
{
    Context &c = ctx();
    int *info = c.einfo.as<int>();
    int refined = 0, done = 0;
    *Vec = H;
    if (c.eig_mode == 1 && c.refine_skip > 0) {
        --c.refine_skip;
    } else if (c.eig_mode == 1) {
        RS_TRY(refine_eigh(H, nb, Vbuf, D, &refined));
        c.eig_refined += refined;
        done = refined;
        if (refined) {
            c.refine_fails = 0;
        } else {
            c.refine_fails = std::min(c.refine_fails + 1, 4);
            c.refine_skip = 1 << (c.refine_fails - 1);
        }
    }
    if (!done && c.eig_mode != 0)
        RS_TRY(eigh_tridiag(H, nb, Vbuf, D, info, &done, c.eig_mode == 2 ? 0 : 1));
    if (done) {
        *Vec = Vbuf;
    } else {
        RS_BLAS(rocsolver_dsyevd(c.blas, rocblas_evect_original, rocblas_fill_upper, nb, H, nb, D,
                                 D + nb, info));
        hipLaunchKernelGGL(k_check_info, dim3(1), dim3(1), 0, c.stream, info, info + 3);
    }
    hipLaunchKernelGGL(k_fix_signs, dim3(nb), dim3(128), 0, c.stream, *Vec, nb);
    RS_HIP(hipGetLastError());
    return kOK;
}

// psi <- psi R^-1 with R^T R = dV psi^T psi, into c.tmp, then swap with c.psi
template <class T>
int gram_schmidt_dev(T *X, int nb)
{
    Context &c = ctx();
    RS_TRY(ensure_dense(nb));
    double *S = c.mat.as<double>();
    RS_TRY(gemm_tn<T>(X, X, nb, c.g.dV, S));
    RS_TRY(chol_inverse(S, nb));
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
    double *D = c.eval.as<double>(), *Vec = nullptr;
    RS_TRY(eigh_sub(H, nb, c.mat3.as<double>(), D, &Vec));
    T *Out = c.tmp.as<T>();
    RS_TRY(gemm_nn<T>(X, Vec, nb, Out));
    RS_HIP(hipMemcpyAsync(esp_dev, D, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
    std::swap(c.psi.p, c.tmp.p);
    std::swap(c.psi.bytes, c.tmp.bytes);
    return kOK;
}

// c.mat <- R^-1, R^T R = dV X^T X
template <class T>
int cholesky_inverse_dev(const T *X, int nb)
{
    Context &c = ctx();
    RS_TRY(ensure_dense(nb));
    double *S = c.mat.as<double>();
    RS_TRY(gemm_tn<T>(X, X, nb, c.g.dV, S));
    return chol_inverse(S, nb);
}

// With Q = X R^-1 (R^-1 in c.mat from cholesky_inverse_dev):
//   Hsub = 0.5 dV (Q^T H Q + (Q^T H Q)^T) = R^-T [0.5 dV (C + C^T)] R^-1,
//   C = X^T HX, then psi <- Q V = X (R^-1 V) in one rotation.
template <class T>
int rayleigh_ritz_factored_dev(T *X, const T *HX, int nb, double *esp_dev)
{
    Context &c = ctx();
    RS_TRY(ensure_dense(nb));
    double *Ri = c.mat.as<double>(), *C = c.mat2.as<double>(), *H = c.mat3.as<double>(),
           *V4 = c.mat4.as<double>();
    const double one = 1.0, zero = 0.0;
    // X^T HX is symmetric (H is) and only 0.5 (C + C^T) is used: the upper
    // tile pairs suffice, the difference is the rounding-level antisymmetric part
    RS_TRY(gemm_tn<T>(X, HX, nb, 1.0, C, true));
    hipLaunchKernelGGL(k_symmetrize, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, C, H, nb,
                       0.5 * c.g.dV);
    RS_HIP(hipGetLastError());
    // C <- Hs R^-1, H <- R^-T C
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, nb, nb, nb, &one,
                          H, nb, Ri, nb, &zero, C, nb));
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, nb, nb, nb,
                          &one, Ri, nb, C, nb, &zero, H, nb));
    double *D = c.eval.as<double>(), *Vec = nullptr;
    RS_TRY(eigh_sub(H, nb, V4, D, &Vec));
    // C <- R^-1 V
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_none, nb, nb, nb, &one,
                          Ri, nb, Vec, nb, &zero, C, nb));
    T *Out = c.tmp.as<T>();
    RS_TRY(gemm_nn<T>(X, C, nb, Out));
    RS_HIP(hipMemcpyAsync(esp_dev, D, sizeof(double) * nb, hipMemcpyDeviceToDevice, c.stream));
    std::swap(c.psi.p, c.tmp.p);
    std::swap(c.psi.bytes, c.tmp.bytes);
    return kOK;
}

// Device time (ms, mean over nrep after one untimed call) of the sweep's tall
// GEMMs on the attached psi (x) and a band-sized work array (y) in the working
// precision: [0] S = dV x^T x, [1] C = x^T y (symmetric product: upper tile
// pairs), [2] out = x U with U = I (nb x nb FP64).
template <class T>
static int time_gemms_t(int nrep, double *ms)
{
    Context &c = ctx();
    const int nb = c.MB;
    RS_TRY(ensure_dense(nb));
    T *X = c.psi.as<T>(), *Y = c.hp.as<T>(), *Out = c.tmp.as<T>();
    double *S = c.mat.as<double>(), *C = c.mat2.as<double>(), *U = c.mat3.as<double>();
    hipLaunchKernelGGL(k_set_identity, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, U, nb);
    RS_HIP(hipGetLastError());
    hipEvent_t a, b;
    RS_HIP(hipEventCreate(&a));
    RS_HIP(hipEventCreate(&b));
    for (int t = 0; t < 3; ++t) {
        for (int rep = -1; rep < nrep; ++rep) {
            if (rep == 0) RS_HIP(hipEventRecord(a, c.stream));
            if (t == 0) RS_TRY(gemm_tn<T>(X, X, nb, c.g.dV, S));
            else if (t == 1) RS_TRY(gemm_tn<T>(X, Y, nb, 1.0, C, true));
            else RS_TRY(gemm_nn<T>(X, U, nb, Out));
        }
        RS_HIP(hipEventRecord(b, c.stream));
        RS_HIP(hipEventSynchronize(b));
        float t_ms = 0.0f;
        RS_HIP(hipEventElapsedTime(&t_ms, a, b));
        ms[t] = nrep > 0 ? t_ms / nrep : 0.0;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return kOK;
}

int time_gemms(int nrep, double *ms)
{
    return ctx().prec == 0 ? time_gemms_t<double>(nrep, ms) : time_gemms_t<float>(nrep, ms);
}

template int convert_from_f64<double>(const double *, double *, long);
template int convert_from_f64<float>(const double *, float *, long);
template int convert_to_f64<double>(const double *, double *, long);
template int convert_to_f64<float>(const float *, double *, long);
template int gram_schmidt_dev<double>(double *, int);
template int gram_schmidt_dev<float>(float *, int);
template int subspace_diag_dev<double>(double *, int, double *);
template int subspace_diag_dev<float>(float *, int, double *);
template int cholesky_inverse_dev<double>(const double *, int);
template int cholesky_inverse_dev<float>(const float *, int);
template int rayleigh_ritz_factored_dev<double>(double *, const double *, int, double *);
template int rayleigh_ritz_factored_dev<float>(float *, const float *, int, double *);

}  // namespace rsdft
