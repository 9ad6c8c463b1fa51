// Symmetric eigensolver of the subspace matrix: Householder reduction to
// tridiagonal form in one cooperative launch, the tridiagonal eigenproblem by
// bisection + inverse iteration (or rocsolver_dstedc), and the
// back-transformation by a blocked wave-per-columns kernel.
//
// The reduction is dsytd2 with the lower convention: A = Q T Q^T,
// Q = H(0) H(1) ... H(n-2), H(i) = I - tau_i v_i v_i^T with v_i(0:i) = 0,
// v_i(i+1) = 1 and v_i(i+2:n) stored in R(i+2:n, i); D, E the diagonal and
// subdiagonal of T (the output format of rocsolver_dsytrd).  At n = 480
// rocsolver_dsytrd takes ~7.5 ms, almost all of it in ~1,500 small launches;
// here the matrix stays in LDS and the reduction takes ~2.5 ms.
//
// nwg workgroups own the columns cyclically (column j in workgroup j % nwg,
// local slot j / nwg) and a column step needs one grid barrier:
//   every workgroup holds v_k, tau_k (computed redundantly, identically),
//   a. p_j = tau_k A(k+1:n, j)^T v_k for its columns j > k (A is symmetric,
//      so this is entry j of tau_k A v_k); it publishes p_j; the owner of
//      column k+1 publishes that column;
//   -- grid barrier --
//   b. every workgroup reads the whole p, forms p^T v_k itself (the same
//      fixed-order sum everywhere), w = p - (tau_k / 2)(p^T v_k) v_k, and
//      A(k+1:n, j) -= v_k w_j + w v_k(j) on its columns; it applies the same
//      update to the published column k+1 and forms v_{k+1}, tau_{k+1} from
//      it (dlarfg).
// The publication area has two slots used by step parity: slot k % 2 is
// written again at step k + 2, after every workgroup has passed the barrier
// of step k + 1 and so has finished reading it.  Published values are
// written and read with agent-scope atomic accesses (coherent across the
// XCDs' L2 caches); every wave waits for its own stores to complete before
// the workgroup signals arrival, so a workgroup that sees all arrival flags
// reads complete data.
#include "common.h"
#include "coop.h"

#include <algorithm>

namespace rsdft {

constexpr int kTrdThreads = 256;
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kTrdMaxWG = 128;

// sum over the workgroup, the same value in every thread; red: kTrdWaves doubles
__device__ __forceinline__ double trd_block_sum(double x, double *red)
{
    x = coop_wave_sum(x);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kTrdWaves; ++i) s += red[i];
    return s;
}

// dlarfg on alpha = col[s+1], x = col[s+2:n]: v = (0.., 1, x / (alpha - beta)),
// beta = -sign(alpha) |(alpha, x)|, tau = (beta - alpha) / beta; tau = 0 and
// beta = alpha when x = 0.  Ends with a barrier.
__device__ void trd_householder(const double *col, double *v, int n, int s, double *red,
                                double &tau, double &beta)
{
    double ss = 0.0;
    for (int i = s + 2 + (int)threadIdx.x; i < n; i += kTrdThreads) ss += col[i] * col[i];
    const double xnorm2 = trd_block_sum(ss, red);
    const double alpha = col[s + 1];
    double scal = 0.0;
    if (xnorm2 == 0.0) {
        tau = 0.0;
        beta = alpha;
    } else {
        beta = -copysign(sqrt(alpha * alpha + xnorm2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
    }
    for (int i = threadIdx.x; i < n; i += kTrdThreads)
        v[i] = i <= s ? 0.0 : (i == s + 1 ? 1.0 : col[i] * scal);
    __syncthreads();
}

// A (n x n, symmetric, full storage) -> D, E, TAU, reflectors in R.
// pub: 4n doubles; flags: 16 nwg words and err zero on entry.
__global__ void __launch_bounds__(kTrdThreads)
k_sytrd_coop(const double *__restrict__ A, int n, int cpw, double *__restrict__ R,
             double *__restrict__ D, double *__restrict__ E, double *__restrict__ TAU,
             double *pub, unsigned *flags, int *err)
{
    extern __shared__ double lds[];
    __shared__ int stop;
    const int nwg = gridDim.x, wg = blockIdx.x, tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    double *cols = lds;                     // [cpw][n] owned columns
    double *v = cols + (size_t)cpw * n;     // [n] reflector of the current step
    double *w = v + n;                      // [n] p, then w
    double *col = w + n;                    // [n] column k+1, updated redundantly
    double *red = col + n;                  // [kTrdWaves]

    for (int c = 0; c < cpw; ++c) {
        const int j = wg + c * nwg;
        if (j >= n) break;
        for (int i = tid; i < n; i += kTrdThreads) cols[(size_t)c * n + i] = A[(size_t)j * n + i];
    }
    for (int i = tid; i < n; i += kTrdThreads) col[i] = A[i];
    if (tid == 0) stop = 0;
    __syncthreads();
    double tau, beta;
    trd_householder(col, v, n, 0, red, tau, beta);
    if (wg == 0) {
        for (int i = 2 + tid; i < n; i += kTrdThreads) R[i] = v[i];
        if (tid == 0) {
            D[0] = col[0];
            E[0] = beta;
            TAU[0] = tau;
        }
    }
    for (int k = 0; k <= n - 3; ++k) {
        double *pp = pub + (size_t)(k & 1) * 2 * n, *pc = pp + n;
        // a. p_j = tau v^T A(:, j) for the owned columns j > k, one wave per column
        for (int c = wave; c < cpw; c += kTrdWaves) {
            const int j = wg + c * nwg;
            if (j >= n) break;
            if (j <= k) continue;
            const double *a = cols + (size_t)c * n;
            double s = 0.0;
            for (int i = k + 1 + lane; i < n; i += 64) s += a[i] * v[i];
            s = tau * coop_wave_sum(s);
            if (lane == 0) coop_put(pp + j, s);
        }
        if ((k + 1) % nwg == wg) {
            const double *a = cols + (size_t)((k + 1) / nwg) * n;
            for (int i = k + 1 + tid; i < n; i += kTrdThreads) coop_put(pc + i, a[i]);
        }
        if (!coop_grid_barrier(flags, (unsigned)(k + 1), err, &stop)) return;
        // b. the whole p and the column k+1; p^T v
        double t = 0.0;
        for (int i = k + 1 + tid; i < n; i += kTrdThreads) {
            const double pi = coop_get(pp + i);
            w[i] = pi;
            col[i] = coop_get(pc + i);
            t += pi * v[i];
        }
        const double alpha = -0.5 * tau * trd_block_sum(t, red);   // ends with a barrier
        for (int i = k + 1 + tid; i < n; i += kTrdThreads) w[i] += alpha * v[i];
        __syncthreads();
        for (int c = 0; c < cpw; ++c) {
            const int j = wg + c * nwg;
            if (j >= n) break;
            if (j <= k) continue;
            const double wj = w[j], vj = v[j];
            double *a = cols + (size_t)c * n;
            for (int i = k + 1 + tid; i < n; i += kTrdThreads) a[i] -= v[i] * wj + w[i] * vj;
        }
        {
            const double wj = w[k + 1], vj = v[k + 1];
            for (int i = k + 1 + tid; i < n; i += kTrdThreads) col[i] -= v[i] * wj + w[i] * vj;
        }
        __syncthreads();
        trd_householder(col, v, n, k + 1, red, tau, beta);
        if ((k + 1) % nwg == wg) {
            for (int i = k + 3 + tid; i < n; i += kTrdThreads) R[(size_t)(k + 1) * n + i] = v[i];
            if (tid == 0) {
                D[k + 1] = col[k + 1];
                E[k + 1] = beta;
                TAU[k + 1] = tau;
            }
        }
    }
    if ((n - 1) % nwg == wg && tid == 0) D[n - 1] = cols[(size_t)((n - 1) / nwg) * n + n - 1];
}

// Z <- Q Z = H(0) H(1) ... H(n-2) Z (dormtr('L', 'L', 'N')).  The columns
// of Z are independent: each wave carries one of them in registers (NPL
// values per lane, n <= 64 NPL).  The reflectors are applied kQBlk at a time
// (i = ib .. ib+kQBlk-1, highest first): with d_i = v_i^T z for the column
// before the block and c_ik = v_i^T v_k,
//     s_i = tau_i (d_i - sum_{k > i} s_k c_ik),   z <- z - sum_i s_i v_i
// is the same as applying them one by one, but the kQBlk dot products share
// one pass over the reflectors and one wave reduction.
constexpr int kQBlk = 8;

__device__ __forceinline__ double trd_v(const double *R, int n, int i, int r)
{
    return r == i + 1 ? 1.0 : (r > i + 1 && r < n ? R[(size_t)i * n + r] : 0.0);
}

// C[b][j][k] = v_{ib+j}^T v_{ib+k} for j < k within block b (one wave each)
__global__ void __launch_bounds__(64)
k_q_overlaps(const double *__restrict__ R, int n, double *__restrict__ C)
{
    const int b = blockIdx.x / (kQBlk * kQBlk), jk = blockIdx.x % (kQBlk * kQBlk);
    const int j = jk / kQBlk, k = jk % kQBlk, lane = threadIdx.x;
    const int i = b * kQBlk + j, i2 = b * kQBlk + k;
    double s = 0.0;
    if (j < k && i2 <= n - 2)
        for (int r = lane; r < n; r += 64) s += trd_v(R, n, i, r) * trd_v(R, n, i2, r);
    s = coop_wave_sum(s);
    if (lane == 0) C[blockIdx.x] = s;
}

// Sum of v[0..K) over the wave, K a power of two <= 64: each exchange at
// distance D keeps half of the values (the half the lane's bit D selects)
// and adds the partner's copy of it; after log2(K) exchanges lane l holds
// value l / (64 / K) summed over its group, the remaining distances finish
// the sum.  Fixed order.
template <int D, int N, int K>
__device__ __forceinline__ void trd_reduce_scatter(double (&v)[K], int lane)
{
    if constexpr (N > 1) {
        constexpr int H = N / 2;
        const bool up = (lane & D) != 0;
#pragma unroll
        for (int i = 0; i < H; ++i) {
            const double send = up ? v[i] : v[i + H];
            const double keep = up ? v[i + H] : v[i];
            v[i] = keep + __shfl_xor(send, D, 64);
        }
        trd_reduce_scatter<D / 2, H, K>(v, lane);
    }
}

// One column per wave, kQWaves columns per workgroup; the workgroup stages
// each block of kQBlk reflectors in LDS (rows padded to 64 NPL), loading the
// next block from global memory while the current one is applied.
constexpr int kQWaves = 4;
template <int NPL>
__global__ void __launch_bounds__(64 * kQWaves)
k_apply_q(const double *__restrict__ R, const double *__restrict__ TAU,
          const double *__restrict__ C, int n, double *Z)
{
    constexpr int NP = 64 * NPL;                         // padded reflector length
    constexpr int NT = 64 * kQWaves;
    constexpr int SPT = kQBlk * NP / NT;                 // staged values per thread
    static_assert(SPT * NT == kQBlk * NP, "whole blocks");
    static_assert(kQBlk <= 64 && (kQBlk & (kQBlk - 1)) == 0, "power of two");
    extern __shared__ __attribute__((aligned(16))) double Vs[];   // [2][kQBlk][NP]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int col = blockIdx.x * kQWaves + wave;
    const bool live = col < n;
    double z[NPL];
#pragma unroll
    for (int m = 0; m < NPL; ++m) {
        const int r = lane + 64 * m;
        z[m] = (live && r < n) ? Z[(size_t)col * n + r] : 0.0;
    }
    const int nblk = (n - 1 + kQBlk - 1) / kQBlk;
    double regs[SPT];
    auto fetch = [&](int b) {
        const int ib = b * kQBlk, nk = min(kQBlk, n - 1 - ib);
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            const int e = tid + NT * j, k = e / NP, r = e % NP;
            regs[j] = k < nk ? trd_v(R, n, ib + k, r) : 0.0;
        }
    };
    auto put = [&](int buf) {
#pragma unroll
        for (int j = 0; j < SPT; ++j) Vs[(size_t)buf * kQBlk * NP + tid + NT * j] = regs[j];
    };
    fetch(nblk - 1);
    put(0);
    __syncthreads();
    for (int bi = 0; bi < nblk; ++bi) {
        const int b = nblk - 1 - bi, cur = bi & 1;
        if (b > 0) fetch(b - 1);
        const double *V = Vs + (size_t)cur * kQBlk * NP;
        if (live) {
            const int ib = b * kQBlk, nk = min(kQBlk, n - 1 - ib);
            double a[kQBlk];
#pragma unroll
            for (int k = 0; k < kQBlk; ++k) {
                double t = 0.0;
#pragma unroll
                for (int m = 0; m < NPL; ++m) t += V[k * NP + lane + 64 * m] * z[m];
                a[k] = t;
            }
            trd_reduce_scatter<32, kQBlk, kQBlk>(a, lane);
            double rsum = a[0];
#pragma unroll
            for (int d = 32 / kQBlk; d > 0; d >>= 1) rsum += __shfl_xor(rsum, d, 64);
            double sk[kQBlk];
#pragma unroll
            for (int k = kQBlk - 1; k >= 0; --k) {
                const double d = __shfl(rsum, k * (64 / kQBlk), 64);
                double t = 0.0;
                if (k < nk) {
                    t = d;
#pragma unroll
                    for (int k2 = k + 1; k2 < kQBlk; ++k2)
                        if (k2 < nk) t -= sk[k2] * C[((size_t)b * kQBlk + k) * kQBlk + k2];
                    t *= TAU[ib + k];
                }
                sk[k] = t;
            }
#pragma unroll
            for (int m = 0; m < NPL; ++m) {
                double t = z[m];
#pragma unroll
                for (int k = 0; k < kQBlk; ++k) t -= sk[k] * V[k * NP + lane + 64 * m];
                z[m] = t;
            }
        }
        if (b > 0) put(cur ^ 1);
        __syncthreads();
    }
    if (live) {
#pragma unroll
        for (int m = 0; m < NPL; ++m) {
            const int r = lane + 64 * m;
            if (r < n) Z[(size_t)col * n + r] = z[m];
        }
    }
}

__global__ void k_trd_clear(unsigned *flags, int nflags, int *err)
{
    for (int i = threadIdx.x; i < nflags; i += blockDim.x) flags[i] = 0u;
    if (threadIdx.x == 0) *err = 0;
}

// flag = -1 when a grid barrier gave up, else the dstedc info
__global__ void k_trd_flag(const int *err, const int *info, int *flag)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (*err != 0) *flag = -1;
        else if (*info != 0) *flag = *info;
    }
}

// ---------------------------------------------------------------------------
// Eigenpairs of the tridiagonal T = tridiag(e, d, e), n <= kTriMax, in place of
// rocsolver_dstedc (whose secular-equation merges take ~2 ms at n = 480):
//  1. eigenvalues: one wave per index j, multisection on Sturm counts (the
//     number of negative pivots of T - sigma I by LAPACK dstebz's recurrence
//     q_i = (d_i - sigma) - e_{i-1}^2 / q_{i-1} with its pivmin guard, the
//     quotient from a Newton-refined hardware reciprocal); each round the 64
//     lanes evaluate kBisS shifts each (independent recurrences, so their
//     latencies overlap), cutting [lo, hi] into 64 kBisS + 1 parts, until
//     hi - lo <= 2 eps max(|lo|, |hi|) + eps |T| / 2;
//  2. eigenvectors: inverse iteration as in LAPACK dstein: LU with partial
//     pivoting of T - lambda_j I (dgttrf, pivots below eps |T| raised to it),
//     kInvRounds solves (dgttrs) from a fixed pseudo-random start, each
//     followed by modified Gram-Schmidt inside clusters of eigenvalues closer
//     than kClusterTol |T| (inverse iteration alone cannot separate those;
//     a second round makes the start vector's component along the wanted
//     eigenvector irrelevant);
//  3. one Newton-Schulz step Z <- Z (3I - Z^T Z) / 2 removes the remaining
//     O(eps |T| / gap) non-orthogonality between vectors of different
//     clusters; it moves vector j by sum_k O(eps |T| / gap_jk) z_k, which
//     changes its residual by O(eps |T|).
// |T| is the Gershgorin bound max(|gl|, |gu|).
// ---------------------------------------------------------------------------
constexpr int kTriMax = 512;
constexpr int kBisS = 2;
constexpr int kInvRounds = 2;
constexpr int kInvVPW = 2;   // eigenvectors per workgroup: (2 x 41 + 16) B x n <= 49 KB of LDS at n = 512
constexpr double kClusterTol = 1e-7;
constexpr double kEps = 1.1102230246251565e-16;   // 2^-53
constexpr double kSafeMin = 2.2250738585072014e-308;

// 1/q: the hardware reciprocal refined by one Newton step (which doubles the
// correct bits); |q| >= pivmin keeps it finite.  A Sturm count from these
// quotients is the exact count of T with each d_i perturbed relatively by
// O(eps), as with IEEE division.
__device__ __forceinline__ double tri_recip(double q)
{
    const double r = __builtin_amdgcn_rcp(q);
    return fma(r, fma(-q, r, 1.0), r);
}

__global__ void __launch_bounds__(64)
k_tri_bisect(const double *__restrict__ d, const double *__restrict__ e, int n,
             double *__restrict__ w, double *__restrict__ tnorm_out, int *__restrict__ info)
{
    __shared__ double sd[kTriMax], se2[kTriMax];
    const int j = blockIdx.x, lane = threadIdx.x;
    double gl = INFINITY, gu = -INFINITY, emax2 = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double el = i > 0 ? fabs(e[i - 1]) : 0.0, er = i < n - 1 ? fabs(e[i]) : 0.0;
        sd[i] = d[i];
        se2[i] = er * er;
        gl = fmin(gl, d[i] - el - er);
        gu = fmax(gu, d[i] + el + er);
        emax2 = fmax(emax2, er * er);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        gl = fmin(gl, __shfl_xor(gl, o, 64));
        gu = fmax(gu, __shfl_xor(gu, o, 64));
        emax2 = fmax(emax2, __shfl_xor(emax2, o, 64));
    }
    __syncthreads();
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double pivmin = kSafeMin * fmax(1.0, emax2);
    double lo = gl - 2.0 * kEps * tnorm - 2.0 * pivmin;
    double hi = gu + 2.0 * kEps * tnorm + 2.0 * pivmin;
    constexpr int K = 64 * kBisS;
    // invariant: count(lo) <= j < count(hi), count(s) = #eigenvalues below s
    for (int round = 0; round < 64; ++round) {
        if (hi - lo <= 2.0 * kEps * fmax(fabs(lo), fabs(hi)) + 0.5 * kEps * tnorm) break;
        const double h = (hi - lo) / (double)(K + 1);
        double q[kBisS], sg[kBisS];
        int cnt[kBisS];
#pragma unroll
        for (int s = 0; s < kBisS; ++s) {
            sg[s] = lo + (double)(lane * kBisS + s + 1) * h;
            double t = sd[0] - sg[s];
            if (fabs(t) < pivmin) t = -pivmin;
            q[s] = t;
            cnt[s] = t < 0.0 ? 1 : 0;
        }
#pragma unroll 8
        for (int i = 1; i < n; ++i) {
            const double di = sd[i], e2 = se2[i - 1];
#pragma unroll
            for (int s = 0; s < kBisS; ++s) {
                double t = (di - sg[s]) - e2 * tri_recip(q[s]);
                if (fabs(t) < pivmin) t = -pivmin;
                q[s] = t;
                cnt[s] += t < 0.0 ? 1 : 0;
            }
        }
        // first shift (in increasing order) with count >= j + 1
        int kl = K;
#pragma unroll
        for (int s = kBisS - 1; s >= 0; --s)
            if (cnt[s] >= j + 1) kl = lane * kBisS + s;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) kl = min(kl, __shfl_xor(kl, o, 64));
        const double nlo = lo + (double)kl * h;                      // shift kl - 1 (or lo)
        const double nhi = kl < K ? lo + (double)(kl + 1) * h : hi;  // shift kl (or hi)
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) {
        w[j] = 0.5 * (lo + hi);
        if (j == 0) {
            *tnorm_out = tnorm;
            *info = 0;
        }
    }
}

__device__ __forceinline__ double tri_start(int i, int j)
{
    // fixed pseudo-random start vector in [-1, 1) (a 32-bit integer mix)
    unsigned h = (unsigned)i * 0x9E3779B1u ^ ((unsigned)j + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)h * (2.0 / 4294967296.0) - 1.0;
}

// X(:, j) <- the fixed pseudo-random start vector, S[j] = 1
__global__ void k_tri_start(double *__restrict__ X, double *__restrict__ S, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= n) return;
    X[(long)i * n + j] = tri_start(i, j);
    if (i == 0) S[j] = 1.0;
}

// one inverse-iteration solve per eigenvector: kInvVPW eigenvectors per
// 64-lane workgroup, lane t < kInvVPW running the sequential recurrences for
// eigenvector j = blockIdx.x * kInvVPW + t with its LU factors (dgttrf: the
// multipliers, reciprocal pivots, first and second superdiagonals and
// row-interchange flags) and right-hand side in LDS ([i][t] layout), so the
// recurrences carry their values in registers and wait only on LDS loads
// that do not depend on them.  The whole wave moves the columns between
// global memory (X(i, j) = X[i*n + j]) and LDS.  A round factors
// T - w_j I and solves with Xin(:, j) * S[j] into Xout, S[j] = 1 / |Xout(:, j)|.
// A workgroup none of whose eigenvectors lies in a cluster (where Gram-Schmidt
// must run between the rounds) does all kInvRounds solves in round 0 with the
// same factors, writes Xfinal (the buffer the last round writes) and skips the
// later rounds.
__global__ void __launch_bounds__(64)
k_tri_invit(const double *__restrict__ d, const double *__restrict__ e,
            const double *__restrict__ w, const double *__restrict__ tnorm_p, int n, int round,
            const double *__restrict__ Xin, double *__restrict__ Xout,
            double *__restrict__ Xfinal, double *__restrict__ S)
{
    extern __shared__ __attribute__((aligned(16))) double tsm[];
    constexpr int V = kInvVPW;
    const int t = threadIdx.x, j0 = blockIdx.x * V;
    const int nv = min(V, n - j0);
    const long N = n;
    const double ctol = kClusterTol * *tnorm_p;
    bool alone = true;   // the same in every thread
    for (int v = 0; v < nv; ++v) {
        const int j = j0 + v;
        if ((j > 0 && w[j] - w[j - 1] < ctol) || (j + 1 < n && w[j + 1] - w[j] < ctol))
            alone = false;
    }
    if (alone && round > 0) return;   // finished in round 0
    const int nsolve = alone ? kInvRounds : 1;
    double *out = alone ? Xfinal : Xout;
    double *FL = tsm, *FR = FL + V * n, *FU = FR + V * n, *FU2 = FU + V * n, *B = FU2 + V * n;
    double *sd = B + V * n, *se = sd + n;
    unsigned char *FP = reinterpret_cast<unsigned char *>(se + n);
    // T and the right-hand sides Xin(:, j) * S[j] into LDS, all lanes
    for (int i = t; i < n; i += 64) {
        sd[i] = d[i];
        se[i] = i < n - 1 ? e[i] : 0.0;
    }
    for (int q = t; q < nv * n; q += 64) {
        const int v = q / n, i = q - v * n;
        B[i * V + v] = Xin[(long)i * N + j0 + v] * S[j0 + v];
    }
    __syncthreads();
    if (t < nv) {
        const int j = j0 + t;
        const double lam = w[j];
        const double ptol = fmax(kEps * *tnorm_p, kSafeMin);
        // dgttrf
        double dc = sd[0] - lam, duc = se[0];   // current D(i), DU(i)
#pragma unroll 8
        for (int i = 0; i < n - 1; ++i) {
            const double li = se[i], dn = sd[i + 1] - lam, dun = se[i + 1];   // se[n-1] = 0
            const bool keep = fabs(dc) >= fabs(li);   // no row interchange
            const double num = keep ? li : dc;
            double den = keep ? dc : li;
            den = den == 0.0 ? 1.0 : den;             // keep with dc = 0 means li = 0
            const double f = num * tri_recip(den);
            double piv = keep ? dc : li;
            const double ndc = keep ? dn - f * duc : duc - f * dn;
            const double nduc = keep ? dun : -f * dun;
            if (fabs(piv) < ptol) piv = piv < 0.0 ? -ptol : ptol;
            FL[i * V + t] = f;
            FR[i * V + t] = tri_recip(piv);
            FU[i * V + t] = keep ? duc : dn;
            FU2[i * V + t] = keep ? 0.0 : dun;
            FP[i * V + t] = keep ? 0 : 1;
            dc = ndc;
            duc = nduc;
        }
        if (fabs(dc) < ptol) dc = dc < 0.0 ? -ptol : ptol;
        FR[(n - 1) * V + t] = tri_recip(dc);
        for (int sv = 0; sv < nsolve; ++sv) {
        if (sv > 0) {   // the next right-hand side: the normalised solution
            const double sc = S[j];
            for (int i = 0; i < n; ++i) B[i * V + t] *= sc;
        }
        // P and L in place: step i writes B(i) and reads B(i+1), not yet written
        double bc = B[t];
#pragma unroll 8
        for (int i = 0; i < n - 1; ++i) {
            const double bn = B[(i + 1) * V + t], f = FL[i * V + t];
            const bool sw = FP[i * V + t] != 0;
            const double y = sw ? bn : bc, o = sw ? bc : bn;
            B[i * V + t] = y;
            bc = o - f * y;
        }
        B[(n - 1) * V + t] = bc;
        // U in place with the sum of squares.  A block of 8 steps grows |x|
        // by at most (2 |T| / ptol)^8 < 1e132, so rescaling by 1e-100 after
        // every block in which |x| passed 1e100 keeps |x| below 1e232.  The
        // block's squares are summed as (x 1e-100)^2 (<= 1e264, finite): the
        // square of the rescaled value, or 1e-200 x^2 without rescale.
        double x1 = 0.0, x2 = 0.0, ysc = 1.0, ss = 0.0;   // x(i+1), x(i+2), scale of the rest of Y
        int i = n - 1;
        while (i >= 0) {
            const int i1 = max(i - 7, 0);
            double mx = 0.0, sb = 0.0;
#pragma unroll 8
            for (int k = 0; k < 8; ++k) {
                const int ii = i - k;
                if (ii < i1) break;
                const double du = ii < n - 1 ? FU[ii * V + t] : 0.0;
                const double du2 = ii < n - 2 ? FU2[ii * V + t] : 0.0;
                const double x = (B[ii * V + t] * ysc - du * x1 - du2 * x2) * FR[ii * V + t];
                B[ii * V + t] = x;
                mx = fmax(mx, fabs(x));
                sb += (x * 1e-100) * (x * 1e-100);
                x2 = x1;
                x1 = x;
            }
            if (mx > 1e100) {   // rescale the computed tail (rare)
                for (int k = i1; k < n; ++k) B[k * V + t] *= 1e-100;
                x1 *= 1e-100;
                x2 *= 1e-100;
                ysc *= 1e-100;
                ss = ss * 1e-200 + sb;
            } else {
                ss += sb * 1e200;   // sb <= 8 here
            }
            i = i1 - 1;
        }
        S[j] = 1.0 / sqrt(ss);
        }
    }
    __syncthreads();
    for (int q = t; q < nv * n; q += 64) {
        const int v = q / n, i = q - v * n;
        out[(long)i * N + j0 + v] = B[i * V + v];
    }
}

// Xout(:, j) *= S[j] for the members of clusters (the others keep the factor),
// then modified Gram-Schmidt inside each cluster (consecutive eigenvalues
// closer than kClusterTol |T|): one wave per cluster, started by its first
// member; members end with S = 1
__global__ void __launch_bounds__(64)
k_tri_mgs(const double *__restrict__ w, const double *__restrict__ tnorm_p, int n,
          double *__restrict__ X, double *__restrict__ S)
{
    const int j = blockIdx.x, lane = threadIdx.x;
    const long N = n;
    const double ctol = kClusterTol * *tnorm_p;
    if (j + 1 >= n || w[j + 1] - w[j] >= ctol) return;
    if (j > 0 && w[j] - w[j - 1] < ctol) return;
    int m = j + 1;
    while (m + 1 < n && w[m + 1] - w[m] < ctol) ++m;
    for (int k = j; k <= m; ++k) {
        const double sk = S[k];
        for (int i = lane; i < n; i += 64) X[i * N + k] *= sk;
    }
    for (int k = j + 1; k <= m; ++k) {
        for (int i0 = j; i0 < k; ++i0) {
            double s = 0.0;
            for (int i = lane; i < n; i += 64) s += X[i * N + i0] * X[i * N + k];
            s = coop_wave_sum(s);
            for (int i = lane; i < n; i += 64) X[i * N + k] -= s * X[i * N + i0];
        }
        double s = 0.0;
        for (int i = lane; i < n; i += 64) s += X[i * N + k] * X[i * N + k];
        const double inv = 1.0 / sqrt(coop_wave_sum(s));
        for (int i = lane; i < n; i += 64) X[i * N + k] *= inv;
    }
    __syncthreads();   // every lane has read S before it is reset
    for (int k = j + lane; k <= m; k += 64) S[k] = 1.0;
}

// X(:, j) *= S[j]
__global__ void k_tri_scale(double *__restrict__ X, const double *__restrict__ S, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j < n) X[(long)i * n + j] *= S[j];
}

__global__ void k_tri_identity(double *G, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) G[(long)j * n + i] = i == j ? 1.0 : 0.0;
}

// eigenvalues of T into D (ascending, replacing the diagonal) and
// eigenvectors into V; scratch: W (2n + 8), X (2 n x n: two iterates), G (n x n)
static int tri_eig(int n, double *D, const double *E, double *V, int *einfo, double *W,
                   double *X, double *G)
{
    Context &c = ctx();
    const size_t nn = (size_t)n * n;
    double *tn = W + n, *S = W + n + 8;
    hipLaunchKernelGGL(k_tri_bisect, dim3(n), dim3(64), 0, c.stream, D, E, n, W, tn, einfo);
    const dim3 ig((n + kInvVPW - 1) / kInvVPW);
    const size_t lds = ((5 * sizeof(double) + 1) * kInvVPW + 2 * sizeof(double)) * (size_t)n;
    double *Xa = X, *Xb = X + nn;
    double *Xfinal = kInvRounds % 2 == 0 ? Xa : Xb;   // where the last round writes
    hipLaunchKernelGGL(k_tri_start, dim3((n + 127) / 128, n), dim3(128), 0, c.stream, Xa, S, n);
    for (int r = 0; r < kInvRounds; ++r) {
        hipLaunchKernelGGL(k_tri_invit, ig, dim3(64), lds, c.stream, D, E, W, tn, n, r, Xa, Xb,
                           Xfinal, S);
        hipLaunchKernelGGL(k_tri_mgs, dim3(n), dim3(64), 0, c.stream, W, tn, n, Xb, S);
        std::swap(Xa, Xb);
    }
    hipLaunchKernelGGL(k_tri_scale, dim3((n + 127) / 128, n), dim3(128), 0, c.stream, Xa, S, n);
    hipLaunchKernelGGL(k_tri_identity, dim3((n + 127) / 128, n), dim3(128), 0, c.stream, G, n);
    RS_HIP(hipGetLastError());
    // Xa holds Z^T column-major: G = 1.5 I - 0.5 Z^T Z = 1.5 I - 0.5 Xa Xa^T, V = Z G = Xa^T G
    const double mh = -0.5, th = 1.5, one = 1.0, zero = 0.0;
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, n,
                          &mh, Xa, n, Xa, n, &th, G, n));
    RS_BLAS(rocblas_dgemm(c.blas, rocblas_operation_transpose, rocblas_operation_none, n, n, n,
                          &one, Xa, n, G, n, &zero, V, n));
    RS_HIP(hipMemcpyAsync(D, W, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
    return kOK;
}

// workgroups and LDS bytes of the cooperative reduction (about 4 columns per
// workgroup), false if n does not fit
static bool trd_shape(int n, int *nwg, int *cpw, size_t *lds)
{
    Context &c = ctx();
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess)
        return false;
    if (n < 3) return false;
    const long budget = (long)(c.max_lds / sizeof(double)) - 3L * n - kTrdWaves - 64;
    if (budget < n) return false;
    const int cmax = (int)std::min<long>(budget / n, 1 << 20);
    int g = std::min(std::min(kTrdMaxWG, cus), (n + 3) / 4);
    g = std::max(g, (n + cmax - 1) / cmax);
    if (g > std::min(kTrdMaxWG, cus)) return false;
    *nwg = g;
    *cpw = (n + g - 1) / g;
    *lds = sizeof(double) * ((size_t)(*cpw) * n + 3 * (size_t)n + kTrdWaves);
    return *lds <= c.max_lds;
}

// Eigenpairs of the symmetric H (n x n, full storage): eigenvalues ascending
// in D, eigenvectors in V.  The tridiagonal eigenproblem goes to tri_eig
// (tri_solver 1, n <= kTriMax) or rocsolver_dstedc.  einfo[0] receives the
// tridiagonal solver's info and einfo[3] a nonzero flag on failure.  Returns
// kOK with *done = 0 when the matrix is too large for the LDS-resident
// reduction (nothing launched then).
int eigh_tridiag(const double *H, int n, double *V, double *D, int *einfo, int *done,
                 int tri_solver)
{
    Context &c = ctx();
    *done = 0;
    int nwg = 0, cpw = 0;
    size_t lds = 0;
    if (!trd_shape(n, &nwg, &cpw, &lds)) return kOK;
    const size_t nn = (size_t)n * n;
    const int nflags = 16 * nwg;
    const int nqb = (n - 1 + kQBlk - 1) / kQBlk;
    const size_t nc = (size_t)nqb * kQBlk * kQBlk;
    const bool own = tri_solver == 1 && n <= kTriMax;
    // tri_eig scratch: W, 2 X, G
    const size_t nown = own ? 3 * nn + 2 * (size_t)n + 8 : 0;
    RS_TRY(c.evwork.ensure(sizeof(double) * (nn + 6 * (size_t)n + nc + nown) +
                           sizeof(int) * (nflags + 4)));
    double *R = c.evwork.as<double>(), *E = R + nn, *TAU = E + n, *pub = TAU + n;
    double *Cq = pub + 4 * (size_t)n;
    double *Wt = Cq + nc, *Xt = Wt + 2 * n + 8, *Gt = Xt + 2 * nn;
    unsigned *flags = reinterpret_cast<unsigned *>(Cq + nc + nown);
    int *err = reinterpret_cast<int *>(flags + nflags);
    if (lds > 64 * 1024)
        RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_sytrd_coop),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_trd_clear, dim3(1), dim3(256), 0, c.stream, flags, nflags, err);
    RS_HIP(hipGetLastError());
    void *args[] = {&H, &n, &cpw, &R, &D, &E, &TAU, &pub, &flags, &err};
    RS_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_sytrd_coop), dim3(nwg),
                                      dim3(kTrdThreads), args, (unsigned)lds, c.stream));
    if (own) RS_TRY(tri_eig(n, D, E, V, einfo, Wt, Xt, Gt));
    else RS_BLAS(rocsolver_dstedc(c.blas, rocblas_evect_tridiagonal, n, D, E, V, n, einfo));
    const dim3 qgrid((n + kQWaves - 1) / kQWaves), qblk(64 * kQWaves);
    const size_t qlds = sizeof(double) * 2 * kQBlk * 64;   // times NPL
    if (n <= 512) {
        if (8 * qlds > 64 * 1024)
            RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_apply_q<8>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)(8 * qlds)));
        if (4 * qlds > 64 * 1024)
            RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_apply_q<4>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)(4 * qlds)));
        hipLaunchKernelGGL(k_q_overlaps, dim3((unsigned)nc), dim3(64), 0, c.stream, R, n, Cq);
    }
    if (n <= 64)
        hipLaunchKernelGGL((k_apply_q<1>), qgrid, qblk, qlds, c.stream, R, TAU, Cq, n, V);
    else if (n <= 128)
        hipLaunchKernelGGL((k_apply_q<2>), qgrid, qblk, 2 * qlds, c.stream, R, TAU, Cq, n, V);
    else if (n <= 256)
        hipLaunchKernelGGL((k_apply_q<4>), qgrid, qblk, 4 * qlds, c.stream, R, TAU, Cq, n, V);
    else if (n <= 512)
        hipLaunchKernelGGL((k_apply_q<8>), qgrid, qblk, 8 * qlds, c.stream, R, TAU, Cq, n, V);
    else
        RS_BLAS(rocsolver_dormtr(c.blas, rocblas_side_left, rocblas_fill_lower,
                                 rocblas_operation_none, n, n, R, n, TAU, V, n));
    hipLaunchKernelGGL(k_trd_flag, dim3(1), dim3(64), 0, c.stream, err, einfo, einfo + 3);
    RS_HIP(hipGetLastError());
    *done = 1;
    return kOK;
}

}  // namespace rsdft
