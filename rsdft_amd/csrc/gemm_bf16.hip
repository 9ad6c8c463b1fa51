// Band-space GEMMs of the FP32-storage path on CDNA4 matrix cores.
//
//   TN:  C(nb x nb) = alpha * A^T B   (overlap S = dV Psi^T Psi, Psi^T H Psi)
//        A, B are (K x nb) column-major with K = ML, so the reduction runs over
//        the grid (373,248 points at Si216) and the output is small: split-K
//        over workgroups, each writing an FP64 partial tile, then a
//        fixed-order FP64 reduction (deterministic).
//   NN:  Out(M x nb) = A(M x nb) * U(nb x nb)   (rotation Psi V, Psi R^-1);
//        mode 1 is a plain FP32 GEMM with a short K and goes to rocBLAS
//        sgemm; mode 2 uses the 64 x 96 tile kernel below
//
// Two arithmetic modes on FP32 inputs:
//   mode 1: v_mfma_f32_16x16x4_f32 products, FP32 accumulation;
//   mode 2: BF16 hi+lo split, x = hi + lo with hi = bf16(x), lo = bf16(x - hi),
//           v_mfma_f32_16x16x32_bf16 products with FP32 accumulation: all four
//           hi/lo products in TN (lo*lo is a positive bias on diagonals such
//           as |psi|^2 if dropped), hi*hi + hi*lo + lo*hi in NN.
// NN is applied in correction form, Out = A diag(U) + A (U - diag(U)): the
// diagonal is an exact FP32 column scale and only the off-diagonal part goes
// through the matrix cores, so the product error scales with the
// off-diagonal of U, which is small for the near-diagonal rotations (R^-1,
// Ritz vectors) of a converging sweep.
// In TN the FP32 accumulators are flushed into FP64 every 128 k, so the
// long K reduction carries FP64 rounding; that keeps S and Hsub accurate to
// ~1e-9 relative instead of the ~1e-5 of one long FP32 chain.
//
// Lane layouts (checked on gfx950): a 16x16xK MFMA takes A[i][k] from lane
// i + 16*(k/kpl) and B[k][j] from lane j + 16*(k/kpl); D[i][j] is register
// i%4 of lane j + 16*(i/4) for the FP32-output products and register i/4 of
// lane j + 16*(i%4) for v_mfma_f64_16x16x4f64.  Any permutation of k shared by A and B leaves
// the product unchanged, so every lane reads 8 consecutive k of its row from
// LDS and the f32 mode walks them as 8 steps of k = 4.
#include "common.h"

#include <algorithm>
#include <type_traits>

namespace rsdft {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int GKB = 32;        // k per LDS stage
constexpr int GLDK = GKB + 4;  // padded LDS row (floats)
constexpr int GFLUSH = 4;      // stages between FP64 flushes in TN

__device__ __forceinline__ long xcd_work(long L, long total) {
    if ((total & 7) != 0) return L;
    const long per = total >> 3;
    return (L & 7) * per + (L >> 3);
}

__device__ __forceinline__ void split_bf16(const float (&x)[8], v8bf &hi, v8bf &lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)x[e];
        hi[e] = h;
        lo[e] = (__bf16)(x[e] - (float)h);
    }
}

__device__ __forceinline__ void lds_read8(const float *p, float (&x)[8]) {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    const float4 b = *reinterpret_cast<const float4 *>(p + 4);
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
    x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
}

// acc[bi][bj] += sum over the KPL k of this lane's A and B rows (KPL a
// multiple of 8: the BF16 products take 8 k per lane and instruction)
template <int MODE, int BB, int KPL>
__device__ __forceinline__ void mfma_block(const float (&af)[BB][KPL], const float (&bf)[BB][KPL],
                                           v4f (&acc)[BB][BB])
{
    if constexpr (MODE == 1) {
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
            for (int bi = 0; bi < BB; ++bi)
#pragma unroll
                for (int bj = 0; bj < BB; ++bj)
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[bi][s], bf[bj][s],
                                                                      acc[bi][bj], 0, 0, 0);
    } else {
        static_assert(KPL % 8 == 0, "8 k per BF16 instruction");
#pragma unroll
        for (int k0 = 0; k0 < KPL; k0 += 8) {
            v8bf ah[BB], al[BB], bh[BB], bl[BB];
#pragma unroll
            for (int b = 0; b < BB; ++b) {
                float ta[8], tb[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    ta[e] = af[b][k0 + e];
                    tb[e] = bf[b][k0 + e];
                }
                split_bf16(ta, ah[b], al[b]);
                split_bf16(tb, bh[b], bl[b]);
            }
#pragma unroll
            for (int bi = 0; bi < BB; ++bi)
#pragma unroll
                for (int bj = 0; bj < BB; ++bj) {
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], bl[bj], acc[bi][bj], 0, 0, 0);
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// TN split-K: partial W[chunk][col*nb + row] over k in [chunk*kc, ...)
// tile pairs (ti, tj) are enumerated from the work index; tile = 32*BB.
// A stage is KB k of the TILE columns of A and B.  Its global loads are
// issued two stages ahead into registers (two register sets), the LDS holds
// two stages (written at the top of a stage, read after its one barrier), so
// the loads of stages s+1 and s+2 are in flight while stage s runs on the
// matrix cores.  The LDS buffer written at stage s was last read at stage
// s-2, before every wave passed the barrier of stage s-1.
// FP32 (MODE 1/2): KB = 32, FP32 accumulators flushed into FP64 every
// GFLUSH stages.  FP64 (MODE 0): KB = 16, v_mfma_f64_16x16x4f64 with FP64
// accumulators, no flush.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

template <class T, int MODE, int BB>
__global__ void __launch_bounds__(256)
k_gemm_tn(const T *__restrict__ A, const T *__restrict__ B, long K, int nb, int nt, int sym,
          int npairs, long kc, int nchunk, double *__restrict__ W)
{
    constexpr bool F64 = sizeof(T) == 8;
    static_assert(F64 == (MODE == 0), "MODE 0 is the FP64 product");
    constexpr int TILE = 32 * BB;
    constexpr int KB = F64 ? 16 : GKB;      // k per stage
    constexpr int LDK = KB + 4;             // padded LDS row (elements)
    constexpr int KPL = KB / 4;             // consecutive k per lane per stage
    constexpr int QR = KB / 4;              // 4-element quads per column per stage
    constexpr int QPT = TILE * QR / 256;    // quads per thread per matrix per stage
    static_assert(QPT * 256 == TILE * QR, "whole quads per thread");
    typedef T Q __attribute__((ext_vector_type(4)));
    typedef typename std::conditional<F64, v4d, v4f>::type Acc;
    __shared__ __attribute__((aligned(32))) T As[2][TILE * LDK];
    __shared__ __attribute__((aligned(32))) T Bs[2][TILE * LDK];
    const long total = (long)gridDim.x;
    const long w = xcd_work(blockIdx.x, total);
    if (w >= (long)npairs * nchunk) return;
    const int chunk = (int)(w / npairs);
    const int p = (int)(w - (long)chunk * npairs);
    int ti, tj;
    if (sym) {   // p enumerates i <= j column by column: p = j(j+1)/2 + i
        tj = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
        while ((tj + 1) * (tj + 2) / 2 <= p) ++tj;
        while (tj * (tj + 1) / 2 > p) --tj;
        ti = p - tj * (tj + 1) / 2;
    } else {
        tj = p / nt;
        ti = p - tj * nt;
    }
    const int i0 = ti * TILE, j0 = tj * TILE;
    const long kbeg = (long)chunk * kc;
    const long kend = min(K, kbeg + kc);
    const long nst = kend > kbeg ? (kend - kbeg + KB - 1) / KB : 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave & 1, wj = wave >> 1;
    const int ar = wi * 16 * BB, bc = wj * 16 * BB;
    const int lrow = lane & 15, lq = lane >> 4;
    const bool vec = (K & 3) == 0;

    Acc acc[BB][BB];
#pragma unroll
    for (int bi = 0; bi < BB; ++bi)
#pragma unroll
        for (int bj = 0; bj < BB; ++bj) acc[bi][bj] = Acc{0, 0, 0, 0};
    double accd[F64 ? 1 : BB][F64 ? 1 : BB][4];
    if constexpr (!F64) {
#pragma unroll
        for (int bi = 0; bi < BB; ++bi)
#pragma unroll
            for (int bj = 0; bj < BB; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) accd[bi][bj][v] = 0.0;
    }

    Q ra[2][QPT], rb[2][QPT];
    // quad it of this thread: column f / QR, k offset 4 (f % QR) of the stage
    auto gload = [&](Q (&qa)[QPT], Q (&qb)[QPT], long kb) {
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int f = tid + it * 256;
            const int col = f / QR, kq = (f - col * QR) * 4;
            const long gk = kb + kq;
            const int ga = i0 + col, gb = j0 + col;
            Q va = Q{0, 0, 0, 0}, vb = Q{0, 0, 0, 0};
            if (vec && gk + 3 < kend) {
                if (ga < nb) va = *reinterpret_cast<const Q *>(A + (long)ga * K + gk);
                if (gb < nb) vb = *reinterpret_cast<const Q *>(B + (long)gb * K + gk);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (gk + e < kend) {
                        if (ga < nb) va[e] = A[(long)ga * K + gk + e];
                        if (gb < nb) vb[e] = B[(long)gb * K + gk + e];
                    }
            }
            qa[it] = va;
            qb[it] = vb;
        }
    };
    auto lstore = [&](const Q (&qa)[QPT], const Q (&qb)[QPT], int buf) {
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int f = tid + it * 256;
            const int col = f / QR, kq = (f - col * QR) * 4;
            *reinterpret_cast<Q *>(&As[buf][col * LDK + kq]) = qa[it];
            *reinterpret_cast<Q *>(&Bs[buf][col * LDK + kq]) = qb[it];
        }
    };
    if (nst > 0) gload(ra[0], rb[0], kbeg);
    if (nst > 1) gload(ra[1], rb[1], kbeg + KB);
    auto stage = [&](auto set_tag, long s, Acc (&ac)[BB][BB]) {
        constexpr int S = decltype(set_tag)::value;
        lstore(ra[S], rb[S], S);
        __syncthreads();
        if (s + 2 < nst) gload(ra[S], rb[S], kbeg + (s + 2) * KB);
        T af[BB][KPL], bf[BB][KPL];
#pragma unroll
        for (int b = 0; b < BB; ++b) {
            const T *pa = &As[S][(ar + b * 16 + lrow) * LDK + KPL * lq];
            const T *pb = &Bs[S][(bc + b * 16 + lrow) * LDK + KPL * lq];
#pragma unroll
            for (int e = 0; e < KPL; e += 4) {
                const Q qa = *reinterpret_cast<const Q *>(pa + e);
                const Q qb = *reinterpret_cast<const Q *>(pb + e);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    af[b][e + v] = qa[v];
                    bf[b][e + v] = qb[v];
                }
            }
        }
        if constexpr (F64) {
#pragma unroll
            for (int st = 0; st < KPL; ++st)
#pragma unroll
                for (int bi = 0; bi < BB; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BB; ++bj)
                        ac[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[bi][st], bf[bj][st],
                                                                         ac[bi][bj], 0, 0, 0);
        } else {
            mfma_block<MODE, BB, KPL>(af, bf, ac);
        }
    };
    // GFLUSH (even) stages per group: the FP32 accumulators start at zero in
    // every group and are added into the FP64 ones at its end, so they stay
    // in the matrix-core registers for the whole group
    static_assert(GFLUSH % 2 == 0, "stage parity repeats per group");
    for (long s = 0; s < nst; s += GFLUSH) {
        if constexpr (F64) {
#pragma unroll
            for (int u = 0; u < GFLUSH; u += 2) {
                if (s + u < nst) stage(std::integral_constant<int, 0>{}, s + u, acc);
                if (s + u + 1 < nst) stage(std::integral_constant<int, 1>{}, s + u + 1, acc);
            }
        } else {
            Acc gacc[BB][BB];
#pragma unroll
            for (int bi = 0; bi < BB; ++bi)
#pragma unroll
                for (int bj = 0; bj < BB; ++bj) gacc[bi][bj] = Acc{0, 0, 0, 0};
#pragma unroll
            for (int u = 0; u < GFLUSH; u += 2) {
                if (s + u < nst) stage(std::integral_constant<int, 0>{}, s + u, gacc);
                if (s + u + 1 < nst) stage(std::integral_constant<int, 1>{}, s + u + 1, gacc);
            }
#pragma unroll
            for (int bi = 0; bi < BB; ++bi)
#pragma unroll
                for (int bj = 0; bj < BB; ++bj)
#pragma unroll
                    for (int v = 0; v < 4; ++v) accd[bi][bj][v] += (double)gacc[bi][bj][v];
        }
    }
    double *Wc = W + (long)chunk * nb * nb;
#pragma unroll
    for (int bi = 0; bi < BB; ++bi)
#pragma unroll
        for (int bj = 0; bj < BB; ++bj) {
            const int col = j0 + bc + bj * 16 + lrow;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // D[i][j] of a 16x16 block is item v of lane j + 16 q with
                // i = 4 q + v (f32 16x16x4) or i = q + 4 v (f64 16x16x4)
                const int row = i0 + ar + bi * 16 + (F64 ? lq + 4 * v : 4 * lq + v);
                if (row < nb && col < nb) {
                    double x;
                    if constexpr (F64) x = acc[bi][bj][v];
                    else x = accd[bi][bj][v];
                    Wc[(long)col * nb + row] = x;
                }
            }
        }
}

// C(i,j) = alpha * sum_chunk W[chunk](i,j); symmetric: tiles below the
// diagonal were not computed, take them from the mirrored element.
__global__ void k_gemm_tn_reduce(const double *__restrict__ W, int nb, int nchunk, double alpha,
                                 int sym, int tile, double *__restrict__ C)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nb) return;
    int si = i, sj = j;
    if (sym && (i / tile) > (j / tile)) { si = j; sj = i; }
    const long off = (long)sj * nb + si, stride = (long)nb * nb;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += W[c * stride + off];
    C[(long)j * nb + i] = alpha * s;
}

// ---------------------------------------------------------------------------
// NN: Out = A * U, A (M x nb) ld M, U (nb x nb) FP32 ld nb.  Tile 64 rows x
// 32*BN columns; wave (wr, wn) owns 32 x 16*BN.  Output staged through LDS
// and stored as float4 runs along the rows.
// ---------------------------------------------------------------------------
// dst = src with a zero diagonal, diag = the diagonal, both FP32
__global__ void k_split_diag_f32(const double *__restrict__ src, float *__restrict__ dst,
                                 float *__restrict__ diag, int n)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)n * n) return;
    const int j = (int)(t / n), i = (int)(t - (long)j * n);
    dst[t] = i == j ? 0.f : (float)src[t];
    if (i == j) diag[j] = (float)src[t];
}

template <int MODE, int BN>
__global__ void __launch_bounds__(256)
k_gemm_nn(const float *__restrict__ A, const float *__restrict__ U, long M, int nb, int ncolb,
          const float *__restrict__ diag, float *__restrict__ Out)
{
    constexpr int TM = 64, TN = 32 * BN;
    constexpr int SA = TM * GLDK, SB = TN * GLDK;
    constexpr int SC = TN * (TM + 4);
    constexpr int SMEM = (SA + SB) > SC ? (SA + SB) : SC;
    __shared__ __attribute__((aligned(16))) float sm[SMEM];
    float *As = sm, *Bs = sm + SA, *Cs = sm;
    const long total = (long)gridDim.x;
    const long nrowb = (M + TM - 1) / TM;
    const long w = xcd_work(blockIdx.x, total);
    if (w >= nrowb * ncolb) return;
    const long rb = w / ncolb;
    const int cb = (int)(w - rb * ncolb);
    const long r0 = rb * TM;
    const int n0 = cb * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave & 1, wn = wave >> 1;
    const int lrow = lane & 15, lq = lane >> 4;
    const bool vec = (M & 3) == 0 && r0 + TM <= M;

    v4f acc[2][BN];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < BN; ++bj) acc[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int kb = 0; kb < nb; kb += GKB) {
        // A block: 32 k x 64 rows, float4 along rows (2 per thread)
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int f = tid + it * 256;
            const int k = f >> 4, rq = (f & 15) * 4;
            const int gk = kb + k;
            float t[4] = {0.f, 0.f, 0.f, 0.f};
            if (gk < nb) {
                const float *src = A + (long)gk * M + r0 + rq;
                if (vec) {
                    const float4 v = *reinterpret_cast<const float4 *>(src);
                    t[0] = v.x; t[1] = v.y; t[2] = v.z; t[3] = v.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (r0 + rq + e < M) t[e] = src[e];
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) As[(rq + e) * GLDK + k] = t[e];
        }
        // U block: TN columns x 32 k, float (U is column-major, k contiguous)
        for (int f = tid; f < TN * GKB; f += 256) {
            const int col = f / GKB, k = f - col * GKB;
            const int gn = n0 + col, gk = kb + k;
            Bs[col * GLDK + k] = (gn < nb && gk < nb) ? U[(long)gn * nb + gk] : 0.f;
        }
        __syncthreads();
        float af[2][8], bf[BN][8];
#pragma unroll
        for (int b = 0; b < 2; ++b) lds_read8(&As[(wr * 32 + b * 16 + lrow) * GLDK + 8 * lq], af[b]);
#pragma unroll
        for (int b = 0; b < BN; ++b) lds_read8(&Bs[(wn * 16 * BN + b * 16 + lrow) * GLDK + 8 * lq], bf[b]);
        __syncthreads();
        if constexpr (MODE == 1) {
#pragma unroll
            for (int s = 0; s < 8; ++s)
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < BN; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[bi][s], bf[bj][s],
                                                                          acc[bi][bj], 0, 0, 0);
        } else {
            v8bf ah[2], al[2], bh[BN], bl[BN];
#pragma unroll
            for (int b = 0; b < 2; ++b) split_bf16(af[b], ah[b], al[b]);
#pragma unroll
            for (int b = 0; b < BN; ++b) split_bf16(bf[b], bh[b], bl[b]);
#pragma unroll
            for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                for (int bj = 0; bj < BN; ++bj) {
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                }
        }
    }
    // stage the 64 x TN tile column-major in LDS, then store float4 runs
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < BN; ++bj) {
            const int c = wn * 16 * BN + bj * 16 + lrow;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = wr * 32 + bi * 16 + 4 * lq + v;
                Cs[c * (TM + 4) + r] = acc[bi][bj][v];
            }
        }
    __syncthreads();
    for (int f = tid; f < TN * (TM / 4); f += 256) {
        const int c = f / (TM / 4), rq = (f - c * (TM / 4)) * 4;
        const int gn = n0 + c;
        if (gn >= nb) continue;
        const float *s = &Cs[c * (TM + 4) + rq];
        float *d = Out + (long)gn * M + r0 + rq;
        const float *a = A + (long)gn * M + r0 + rq;
        if (vec) {
            float4 o = make_float4(s[0], s[1], s[2], s[3]);
            if (diag) {
                const float dg = diag[gn];
                const float4 av = *reinterpret_cast<const float4 *>(a);
                o.x += dg * av.x; o.y += dg * av.y; o.z += dg * av.z; o.w += dg * av.w;
            }
            *reinterpret_cast<float4 *>(d) = o;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (r0 + rq + e < M) d[e] = s[e] + (diag ? diag[gn] * a[e] : 0.f);
        }
    }
}

__global__ void k_to_f32(const double *src, float *dst, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
// sym_result: A^T B is known to be symmetric (A == B, or B = H A with H
// symmetric): only the tile pairs i <= j are computed, the rest mirrored
template <class T, int MODE>
static int gemm_tn_mode(const T *A, const T *B, int nb, long K, double alpha, double *C,
                        bool sym_result)
{
    Context &c = ctx();
    constexpr int BB = 2, TILE = 32 * BB;
    constexpr int KB = sizeof(T) == 8 ? 16 : GKB;
    const int nt = (nb + TILE - 1) / TILE;
    const bool sym = A == B || sym_result;
    const int npairs = sym ? nt * (nt + 1) / 2 : nt * nt;
    // enough workgroups to fill 256 CUs several times over
    int nchunk = (int)std::max<long>(1, std::min<long>((4096 + npairs - 1) / npairs, (K + 4095) / 4096));
    long kc = (K + nchunk - 1) / nchunk;
    kc = (kc + KB - 1) / KB * KB;
    nchunk = (int)((K + kc - 1) / kc);
    RS_TRY(c.gemm_part.ensure(sizeof(double) * (size_t)nchunk * nb * nb));
    double *W = c.gemm_part.as<double>();
    long nwork = (long)npairs * nchunk;
    long grid = (nwork + 7) / 8 * 8;
    hipLaunchKernelGGL((k_gemm_tn<T, MODE, BB>), dim3(grid), dim3(256), 0, c.stream, A, B, K, nb,
                       nt, sym ? 1 : 0, npairs, kc, nchunk, W);
    RS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gemm_tn_reduce, dim3((nb + 127) / 128, nb), dim3(128), 0, c.stream, W, nb,
                       nchunk, alpha, sym ? 1 : 0, TILE, C);
    RS_HIP(hipGetLastError());
    return kOK;
}

// FP64 A^T B on the FP64 matrix cores (same split-K tiling and fixed-order
// reduction as the FP32 products)
int mfma_gemm_tn_f64(const double *A, const double *B, int nb, long K, double alpha, double *C,
                     int sym_result)
{
    return gemm_tn_mode<double, 0>(A, B, nb, K, alpha, C, sym_result != 0);
}

// mode 1: the rotation is a plain FP32 GEMM (K = nb is short), rocBLAS sgemm
static int gemm_nn_sgemm(const float *A, const double *U, int nb, long M, float *Out)
{
    Context &c = ctx();
    RS_TRY(c.matT.ensure(sizeof(float) * ((size_t)nb * nb + nb)));
    float *U32 = c.matT.as<float>();
    const long nn = (long)nb * nb;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::min<long>((nn + 255) / 256, 1024)), dim3(256),
                       0, c.stream, U, U32, nn);
    RS_HIP(hipGetLastError());
    const float one = 1.0f, zero = 0.0f;
    RS_BLAS(rocblas_sgemm(c.blas, rocblas_operation_none, rocblas_operation_none, (int)M, nb, nb,
                          &one, A, (int)M, U32, nb, &zero, Out, (int)M));
    return kOK;
}

template <int MODE>
static int gemm_nn_mode(const float *A, const double *U, int nb, long M, float *Out)
{
    Context &c = ctx();
    if constexpr (MODE == 1) return gemm_nn_sgemm(A, U, nb, M, Out);
    constexpr int BN = 3, TN = 32 * BN;
    RS_TRY(c.matT.ensure(sizeof(float) * ((size_t)nb * nb + nb)));
    float *U32 = c.matT.as<float>();
    float *D32 = U32 + (size_t)nb * nb;
    const long nn = (long)nb * nb;
    hipLaunchKernelGGL(k_split_diag_f32, dim3((nn + 255) / 256), dim3(256), 0, c.stream, U, U32, D32, nb);
    const int ncolb = (nb + TN - 1) / TN;
    const long nrowb = (M + 63) / 64;
    const long nwork = nrowb * ncolb;
    const long grid = (nwork + 7) / 8 * 8;
    hipLaunchKernelGGL((k_gemm_nn<MODE, BN>), dim3(grid), dim3(256), 0, c.stream, A, U32, M, nb,
                       ncolb, D32, Out);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int split_gemm_tn(const float *A, const float *B, int nb, long K, double alpha, double *C,
                  int sym_result)
{
    return ctx().prec == 2 ? gemm_tn_mode<float, 2>(A, B, nb, K, alpha, C, sym_result != 0)
                           : gemm_tn_mode<float, 1>(A, B, nb, K, alpha, C, sym_result != 0);
}

// fp32 != 0: FP32 products whatever the precision mode
template <class T>
int split_gemm_nn(const float *A, const double *U, int nb, long M, float *Out, int fp32)
{
    return ctx().prec == 2 && !fp32 ? gemm_nn_mode<2>(A, U, nb, M, Out)
                                    : gemm_nn_mode<1>(A, U, nb, M, Out);
}

template int split_gemm_tn<float>(const float *, const float *, int, long, double, double *, int);
template int split_gemm_tn<double>(const float *, const float *, int, long, double, double *, int);
template int split_gemm_nn<float>(const float *, const double *, int, long, float *, int);
template int split_gemm_nn<double>(const float *, const double *, int, long, float *, int);

}  // namespace rsdft
