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
//   mode 2: BF16 three-piece split, x = hi + mid + lo (each a BF16 holding the
//           next 8 bits of x), v_mfma_f32_16x16x32_bf16 products with FP32
//           accumulation: the six piece products of order >= 2^-16 (hi*hi,
//           hi*mid, mid*hi, hi*lo, mid*mid, lo*hi), so each product is good
//           to ~2^-25 relative, the accuracy of the FP32 products of mode 1,
//           at six BF16 instructions per 32 k where mode 1 takes eight FP32
//           instructions of 16x the cycles per flop.
// (The FP64 path's TN products use the same split-K kernel with
// v_mfma_f64_16x16x4f64 and FP64 accumulators, MODE 0.)
// Mode 2 applies NN in correction form, Out = A diag(U) + A (U - diag(U)): the
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
constexpr int GFLUSH = 4;      // stages between FP64 flushes in TN

__device__ __forceinline__ long xcd_work(long L, long total) {
    if ((total & 7) != 0) return L;
    const long per = total >> 3;
    return (L & 7) * per + (L >> 3);
}

// x = hi + mid + lo with hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid):
// both differences are exact in FP32, so the three pieces carry all 24 bits
// of x (the residual left is below 2^-26 |x|).  The kernels below split each
// element once, where it is staged into LDS.

// acc += a * b over one 16x16x32 block from the three-piece splits: the six
// products down to order 2^-16 relative, smallest first; the dropped
// mid*lo, lo*mid and lo*lo terms are below 2^-25 of |a||b|
__device__ __forceinline__ v4f mfma_split3(const v8bf &ah, const v8bf &am, const v8bf &al,
                                           const v8bf &bh, const v8bf &bm, const v8bf &bl, v4f acc)
{
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
}

// acc[bi][bj] += sum over the KPL k of this lane's A and B rows, FP32
// products (the lane's KPL consecutive k walked as KPL steps of k = 4)
template <int BB, int KPL>
__device__ __forceinline__ void mfma_block_f32(const float (&af)[BB][KPL],
                                               const float (&bf)[BB][KPL], v4f (&acc)[BB][BB])
{
#pragma unroll
    for (int s = 0; s < KPL; ++s)
#pragma unroll
        for (int bi = 0; bi < BB; ++bi)
#pragma unroll
            for (int bj = 0; bj < BB; ++bj)
                acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[bi][s], bf[bj][s],
                                                                  acc[bi][bj], 0, 0, 0);
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
    static_assert(MODE == 0 || MODE == 1, "the BF16 split products run in k_gemm_tn3");
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
        {
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
                            ac[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                                af[bi][st], bf[bj][st], ac[bi][bj], 0, 0, 0);
            } else {
                mfma_block_f32<BB, KPL>(af, bf, ac);
            }
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

// ---------------------------------------------------------------------------
// TN split-K in the three-piece BF16 mode on 128 x 128 tiles: 512 threads,
// wave (wi, wj) = (w % 4, w / 4) owns rows 32 wi .. 32 wi + 31 and columns
// 64 wj .. 64 wj + 63 of the tile (2 x 4 blocks of 16 x 16).  The six piece
// products per block make the matrix cores fast enough that the operand
// stream from L2 bounds the 64-wide tiles; doubling the tile halves the
// operand bytes per product.  Global loads run two stages ahead in registers,
// two LDS stages of split pieces, one barrier per stage (order below), FP64
// flush every GFLUSH stages as in k_gemm_tn; the two LDS stages (120 KB) are
// dynamic shared memory.
// ---------------------------------------------------------------------------
constexpr int kTn3Tile = 128;
constexpr int kTn3Threads = 512;
constexpr int kTn3Ldk = GKB;       // bf16 pitch of a piece row (64 bytes, chunks swizzled)

// LDS position of the 16-byte chunk c (k = 8c .. 8c+7) of row r of a piece:
// chunk c ^ ((r >> 2) & 3).  Rows are 64 bytes, so a store of 8 rows x 4
// chunks fills the 64 banks exactly twice (the least for 512 bytes), and the
// 16 lanes that read one chunk of 16 consecutive rows meet 16 different
// 4-bank groups.
__device__ __forceinline__ int tn3_swz(int r, int c) { return c ^ ((r >> 2) & 3); }
constexpr size_t kTn3Stage = 2 * 3 * (size_t)kTn3Tile * kTn3Ldk * sizeof(__bf16);

__global__ void __launch_bounds__(kTn3Threads)
k_gemm_tn3(const float *__restrict__ A, const float *__restrict__ B, long K, int nb, int nt,
           int sym, int npairs, long kc, int nchunk, double *__restrict__ W)
{
    constexpr int TILE = kTn3Tile, KB = GKB, LDKB = kTn3Ldk, PIECE = TILE * LDKB;
    constexpr int QR = KB / 4;                      // 4-element quads per column per stage
    constexpr int QPT = TILE * QR / kTn3Threads;    // quads per thread per matrix per stage
    static_assert(QPT * kTn3Threads == TILE * QR, "whole quads per thread");
    constexpr int BI = 2, BJ = 4;
    typedef float Q __attribute__((ext_vector_type(4)));
    typedef __bf16 v4bf __attribute__((ext_vector_type(4)));
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_tn3[];
    auto piece = [&](int b, int m, int q) {
        return reinterpret_cast<__bf16 *>(lds_tn3 + b * kTn3Stage) + (m * 3 + q) * PIECE;
    };
    const long w = xcd_work(blockIdx.x, (long)gridDim.x);
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
    const int ar = (wave & 3) * 32, bc = (wave >> 2) * 64;
    const int lrow = lane & 15, lq = lane >> 4;
    const bool vec = (K & 3) == 0;
    // the lane's chunk lq of rows 16 m + lrow: ar, bc and 16 b are multiples
    // of 16, so tn3_swz depends on lrow only
    const int rchunk = 8 * tn3_swz(lrow, lq);

    double accd[BI][BJ][4];
#pragma unroll
    for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < BJ; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) accd[bi][bj][v] = 0.0;

    Q ra[2][QPT], rb[2][QPT];
    auto gload = [&](Q (&qa)[QPT], Q (&qb)[QPT], long kb) {
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int f = tid + it * kTn3Threads;
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
            const int f = tid + it * kTn3Threads;
            const int col = f / QR, kq = (f - col * QR) * 4;
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const Q &x = m == 0 ? qa[it] : qb[it];
                v4bf h, md, lo;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const __bf16 hh = (__bf16)x[e];
                    const float r = x[e] - (float)hh;
                    const __bf16 mm = (__bf16)r;
                    h[e] = hh;
                    md[e] = mm;
                    lo[e] = (__bf16)(r - (float)mm);
                }
                const int at = col * LDKB + 8 * tn3_swz(col, kq >> 3) + (kq & 7);
                *reinterpret_cast<v4bf *>(piece(buf, m, 0) + at) = h;
                *reinterpret_cast<v4bf *>(piece(buf, m, 1) + at) = md;
                *reinterpret_cast<v4bf *>(piece(buf, m, 2) + at) = lo;
            }
        }
    };
    // Stage s: after the barrier every wave reads its operands of stage s from
    // LDS buffer s % 2, then splits and stores stage s + 1 (register set
    // (s + 1) % 2) into the other buffer and refills that set with stage
    // s + 3, while its matrix-core products of stage s run: the VALU split
    // overlaps the MFMA work of the same wave.  The buffer written at stage s
    // was last read at stage s - 1, before every wave passed this barrier.
    if (nst > 0) gload(ra[0], rb[0], kbeg);
    if (nst > 1) gload(ra[1], rb[1], kbeg + KB);
    if (nst > 0) {
        lstore(ra[0], rb[0], 0);
        if (nst > 2) gload(ra[0], rb[0], kbeg + 2 * KB);
    }
    auto stage = [&](auto set_tag, long s, v4f (&ac)[BI][BJ]) {
        constexpr int S = decltype(set_tag)::value;   // s % 2
        __syncthreads();
        v8bf pa[3][BI], pb[3][BJ];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int b = 0; b < BI; ++b) {
                pa[q][b] = *reinterpret_cast<const v8bf *>(piece(S, 0, q) +
                                                           (ar + b * 16 + lrow) * LDKB + rchunk);
            }
#pragma unroll
            for (int b = 0; b < BJ; ++b) {
                pb[q][b] = *reinterpret_cast<const v8bf *>(piece(S, 1, q) +
                                                           (bc + b * 16 + lrow) * LDKB + rchunk);
            }
        }
        if (s + 1 < nst) {
            lstore(ra[1 - S], rb[1 - S], 1 - S);
            if (s + 3 < nst) gload(ra[1 - S], rb[1 - S], kbeg + (s + 3) * KB);
        }
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
            for (int bj = 0; bj < BJ; ++bj)
                ac[bi][bj] = mfma_split3(pa[0][bi], pa[1][bi], pa[2][bi], pb[0][bj], pb[1][bj],
                                         pb[2][bj], ac[bi][bj]);
    };
    static_assert(GFLUSH % 2 == 0, "stage parity repeats per group");
    for (long s = 0; s < nst; s += GFLUSH) {
        v4f gacc[BI][BJ];
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
            for (int bj = 0; bj < BJ; ++bj) gacc[bi][bj] = v4f{0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < GFLUSH; u += 2) {
            if (s + u < nst) stage(std::integral_constant<int, 0>{}, s + u, gacc);
            if (s + u + 1 < nst) stage(std::integral_constant<int, 1>{}, s + u + 1, gacc);
        }
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
            for (int bj = 0; bj < BJ; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) accd[bi][bj][v] += (double)gacc[bi][bj][v];
    }
    double *Wc = W + (long)chunk * nb * nb;
#pragma unroll
    for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < BJ; ++bj) {
            const int col = j0 + bc + bj * 16 + lrow;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = i0 + ar + bi * 16 + 4 * lq + v;
                if (row < nb && col < nb) Wc[(long)col * nb + row] = accd[bi][bj][v];
            }
        }
}

// C(i,j) = alpha * sum_chunk W[chunk](i,j), summed in chunk order, one
// element per thread.  A workgroup covers 32 rows x 8 columns of C and reads
// W along the columns (coalesced).  Symmetric products: the GEMM computed
// only the tile pairs ti <= tj; a block inside such a pair also writes its
// transpose when the pair is off the diagonal (through LDS, as 64-byte runs),
// and the blocks below the diagonal tiles return.
__global__ void __launch_bounds__(256)
k_gemm_tn_reduce(const double *__restrict__ W, int nb, int nchunk, double alpha, int sym,
                 int tile, double *__restrict__ C)
{
    __shared__ double tr[8][33];
    const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 8;
    if (sym && i0 / tile > j0 / tile) return;
    const bool mirror = sym && i0 / tile < j0 / tile;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int i = i0 + tx, j = j0 + ty;
    double s = 0.0;
    if (i < nb && j < nb) {
        const long off = (long)j * nb + i, stride = (long)nb * nb;
        for (int c = 0; c < nchunk; ++c) s += W[c * stride + off];
        C[(long)j * nb + i] = alpha * s;
    }
    if (!mirror) return;
    tr[ty][tx] = alpha * s;
    __syncthreads();
    const int a = threadIdx.x & 7, b = threadIdx.x >> 3;   // C(j0 + a, i0 + b) = C(i0 + b, j0 + a)
    if (j0 + a < nb && i0 + b < nb) C[(long)(i0 + b) * nb + j0 + a] = tr[a][b];
}

// ---------------------------------------------------------------------------
// NN (BF16 three-piece mode): Out = A * U, A (M x nb) ld M, U (nb x nb) FP64
// ld nb, in correction form Out = A diag(U) + A E, E = U - diag(U).
// k_split_u3 rounds E to FP32 and splits it once into its three BF16 pieces,
// [col][k] with k padded to a multiple of 32 (zeros), so the GEMM copies them
// into LDS without arithmetic.  Tile kNnRows rows x 32*BN columns, one wave
// per 32 rows x 16*BN columns.  Every k step the loading thread splits its
// 2 x 4 elements of A once and stores the pieces [row][k] (one 32-bit store
// per row and piece: two consecutive k); the lanes then read 8 consecutive k
// of their row or column per piece, the v_mfma_f32_16x16x32_bf16 operand
// layout.  The next step's A and U are loaded into registers while the
// current step is read and multiplied.  The output is staged through LDS and
// stored as float4 runs along the rows, with the diagonal term A diag(U)
// added in FP32.
// ---------------------------------------------------------------------------
constexpr int kNnRows = 64;                    // rows of A per NN tile
constexpr int kNnThreads = 2 * 64 * (kNnRows / 32);   // row groups of 32 x 2 column groups

__global__ void k_split_u3(const double *__restrict__ U, int nb, int kpad,
                           __bf16 *__restrict__ piece, float *__restrict__ diag)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const long n = (long)nb * kpad;
    if (t >= n) return;
    const int col = (int)(t / kpad), k = (int)(t - (long)col * kpad);
    float x = 0.f;
    if (k < nb) {
        const float u = (float)U[(long)col * nb + k];
        if (k == col) diag[col] = u;
        else x = u;
    }
    const __bf16 h = (__bf16)x;
    const float r = x - (float)h;
    const __bf16 m = (__bf16)r;
    piece[t] = h;
    piece[n + t] = m;
    piece[2 * n + t] = (__bf16)(r - (float)m);
}

// LDS position of the 16-byte chunk c (k = 8c .. 8c+7) of row r of an A
// piece: chunk c ^ ((r >> 4) & 3).  The loading threads store rows 4 apart
// (80-byte row pitch): unswizzled, rows 16 apart land on the same banks (a
// 4-way conflict on every store); with the swizzle the 64 lanes of a store
// cover the 64 banks once, and the 16 lanes that read one chunk of 16
// consecutive rows still do.  The U pieces, copied as whole 16-byte chunks,
// use the TN layout (64-byte rows, tn3_swz).
__device__ __forceinline__ int nn_swz(int r, int c) { return c ^ ((r >> 4) & 3); }

template <int BN>
__global__ void __launch_bounds__(kNnThreads)
k_gemm_nn(const float *__restrict__ A, const __bf16 *__restrict__ Up, int kpad, long M, int nb,
          int ncolb, const float *__restrict__ diag, float *__restrict__ Out)
{
    constexpr int TM = kNnRows, TN = 32 * BN, NT = kNnThreads;
    constexpr int LDKB = GKB + 8;                      // bf16 pitch of a piece row
    constexpr int LDKU = GKB;                          // U pieces: 64-byte rows, tn3_swz chunks
    constexpr int PA = TM * LDKB, PB = TN * LDKU;      // elements per piece array
    constexpr int SIN = 3 * (PA + PB) * 2;             // bytes of the staged pieces
    constexpr int SC = TN * (TM + 4) * 4;              // bytes of the output staging
    constexpr int SMEM = SIN > SC ? SIN : SC;
    constexpr int NU = 3 * TN * 4;                     // 16-byte U piece copies per k step
    constexpr int UPT = (NU + NT - 1) / NT;            // per thread
    static_assert(TM / 4 * (GKB / 2) == NT, "one row quad x k pair of A per thread");
    __shared__ __attribute__((aligned(16))) unsigned char sm[SMEM];
    __bf16 *Ap = reinterpret_cast<__bf16 *>(sm);       // piece q at Ap + q * PA
    __bf16 *Bp = Ap + 3 * PA;                          // piece q at Bp + q * PB
    float *Cs = reinterpret_cast<float *>(sm);
    const long total = (long)gridDim.x;
    const long nrowb = (M + TM - 1) / TM;
    const long w = xcd_work(blockIdx.x, total);
    if (w >= nrowb * ncolb) return;
    const long rb = w / ncolb;
    const int cb = (int)(w - rb * ncolb);
    const long r0 = rb * TM;
    const int n0 = cb * TN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave % (TM / 32), wn = wave / (TM / 32);   // rows 32 wr, columns 16 BN wn
    const int lrow = lane & 15, lq = lane >> 4;
    const bool vec = (M & 3) == 0 && r0 + TM <= M;
    const long pn = (long)nb * kpad;                   // elements per piece of U
    // A loads: thread t owns rows rq .. rq+3 at k = kb + kp, kb + kp + 1
    const int rq = (tid % (TM / 4)) * 4, kp = (tid / (TM / 4)) * 2;
    // the lane's U chunk lq of column 16 m + lrow (wn 16 BN and 16 b are
    // multiples of 16, so tn3_swz depends on lrow only)
    const int uchunk = 8 * tn3_swz(lrow, lq);

    v4f acc[2][BN];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < BN; ++bj) acc[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};

    float ta[2][4];
    uint4 tu[UPT];
    // global operands of the k step starting at kb into registers
    auto gload = [&](int kb) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int gk = kb + kp + j;
#pragma unroll
            for (int e = 0; e < 4; ++e) ta[j][e] = 0.f;
            if (gk < nb) {
                const float *src = A + (long)gk * M + r0 + rq;
                if (vec) {
                    const float4 v = *reinterpret_cast<const float4 *>(src);
                    ta[j][0] = v.x; ta[j][1] = v.y; ta[j][2] = v.z; ta[j][3] = v.w;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (r0 + rq + e < M) ta[j][e] = src[e];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UPT; ++u) {
            const int f = tid + u * NT;
            tu[u] = make_uint4(0u, 0u, 0u, 0u);
            if (f < NU) {
                const int q = f / (TN * 4), rem = f - q * (TN * 4);
                const int col = rem >> 2, c8 = (rem & 3) * 8;
                const int gn = n0 + col;
                if (gn < nb)
                    tu[u] = *reinterpret_cast<const uint4 *>(Up + q * pn + (long)gn * kpad + kb + c8);
            }
        }
    };
    gload(0);
    for (int kb = 0; kb < nb; kb += GKB) {
        // split the staged A elements (one 32-bit store per row and piece:
        // two consecutive k) and store the U pieces
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            __bf16 pc[3][2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const float x = ta[j][e];
                const __bf16 h = (__bf16)x;
                const float r = x - (float)h;
                const __bf16 m = (__bf16)r;
                pc[0][j] = h;
                pc[1][j] = m;
                pc[2][j] = (__bf16)(r - (float)m);
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
                const int row = rq + e;
                *reinterpret_cast<v2bf *>(Ap + q * PA + row * LDKB + 8 * nn_swz(row, kp >> 3) +
                                          (kp & 7)) = v2bf{pc[q][0], pc[q][1]};
            }
        }
#pragma unroll
        for (int u = 0; u < UPT; ++u) {
            const int f = tid + u * NT;
            if (f < NU) {
                const int q = f / (TN * 4), rem = f - q * (TN * 4);
                const int col = rem >> 2, c8 = (rem & 3) * 8;
                *reinterpret_cast<uint4 *>(Bp + q * PB + col * LDKU + 8 * tn3_swz(col, c8 >> 3)) =
                    tu[u];
            }
        }
        __syncthreads();
        // the next step's operands are in flight while this step's are read
        // and multiplied
        if (kb + GKB < nb) gload(kb + GKB);
        v8bf pa[3][2], pb[3][BN];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int row = wr * 32 + b * 16 + lrow;
                pa[q][b] = *reinterpret_cast<const v8bf *>(Ap + q * PA + row * LDKB +
                                                           8 * nn_swz(row, lq));
            }
#pragma unroll
            for (int b = 0; b < BN; ++b)
                pb[q][b] = *reinterpret_cast<const v8bf *>(
                    Bp + q * PB + (wn * 16 * BN + b * 16 + lrow) * LDKU + uchunk);
        }
        __syncthreads();
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < BN; ++bj)
                acc[bi][bj] = mfma_split3(pa[0][bi], pa[1][bi], pa[2][bi], pb[0][bj], pb[1][bj],
                                          pb[2][bj], acc[bi][bj]);
    }
    // stage the TM x TN tile column-major in LDS, then store float4 runs
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
    for (int f = tid; f < TN * (TM / 4); f += NT) {
        const int c = f / (TM / 4), rq4 = (f - c * (TM / 4)) * 4;
        const int gn = n0 + c;
        if (gn >= nb) continue;
        const float *s = &Cs[c * (TM + 4) + rq4];
        float *d = Out + (long)gn * M + r0 + rq4;
        const float *a = A + (long)gn * M + r0 + rq4;
        const float dg = diag[gn];
        if (vec) {
            const float4 av = *reinterpret_cast<const float4 *>(a);
            const float4 o = make_float4(s[0] + dg * av.x, s[1] + dg * av.y, s[2] + dg * av.z,
                                         s[3] + dg * av.w);
            *reinterpret_cast<float4 *>(d) = o;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (r0 + rq4 + e < M) d[e] = s[e] + dg * a[e];
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
// three-piece BF16 mode: the 128-wide tile kernel
static int gemm_tn_split3(const float *A, const float *B, int nb, long K, double alpha,
                          double *C, bool sym_result)
{
    Context &c = ctx();
    hipStream_t st = c.stream;
    constexpr int TILE = kTn3Tile;
    const int nt = (nb + TILE - 1) / TILE;
    const bool sym = A == B || sym_result;
    const int npairs = sym ? nt * (nt + 1) / 2 : nt * nt;
    // about five rounds of one 512-thread workgroup per CU
    int nchunk = (int)std::max<long>(1, std::min<long>((1280 + npairs - 1) / npairs, (K + 1023) / 1024));
    long kc = (K + nchunk - 1) / nchunk;
    kc = (kc + GKB - 1) / GKB * GKB;
    nchunk = (int)((K + kc - 1) / kc);
    RS_TRY(c.gemm_part.ensure(sizeof(double) * (size_t)nchunk * nb * nb));
    double *W = c.gemm_part.as<double>();
    static bool lds_set = false;
    if (!lds_set) {
        RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k_gemm_tn3),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(2 * kTn3Stage)));
        lds_set = true;
    }
    const long nwork = (long)npairs * nchunk;
    const long grid = (nwork + 7) / 8 * 8;
    hipLaunchKernelGGL(k_gemm_tn3, dim3(grid), dim3(kTn3Threads), 2 * kTn3Stage, st, A, B,
                       K, nb, nt, sym ? 1 : 0, npairs, kc, nchunk, W);
    RS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gemm_tn_reduce, dim3((nb + 31) / 32, (nb + 7) / 8), dim3(256), 0, st, W, nb,
                       nchunk, alpha, sym ? 1 : 0, TILE, C);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T, int MODE>
static int gemm_tn_mode(const T *A, const T *B, int nb, long K, double alpha, double *C,
                        bool sym_result)
{
    Context &c = ctx();
    hipStream_t st = c.stream;
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
    hipLaunchKernelGGL((k_gemm_tn<T, MODE, BB>), dim3(grid), dim3(256), 0, st, A, B, K, nb,
                       nt, sym ? 1 : 0, npairs, kc, nchunk, W);
    RS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gemm_tn_reduce, dim3((nb + 31) / 32, (nb + 7) / 8), dim3(256), 0, st, W, nb,
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

// mode 2: the three-piece BF16 kernel above
static int gemm_nn_split3(const float *A, const double *U, int nb, long M, float *Out)
{
    Context &c = ctx();
    constexpr int BN = 3, TN = 32 * BN;
    const int kpad = (nb + GKB - 1) / GKB * GKB;
    const long pn = (long)nb * kpad;
    RS_TRY(c.matT.ensure(sizeof(__bf16) * 3 * (size_t)pn + sizeof(float) * (size_t)nb + 16));
    __bf16 *Up = c.matT.as<__bf16>();
    float *D32 = reinterpret_cast<float *>(Up + 3 * pn + 8 - (3 * pn) % 8);
    hipLaunchKernelGGL(k_split_u3, dim3((unsigned)((pn + 255) / 256)), dim3(256), 0, c.stream, U,
                       nb, kpad, Up, D32);
    RS_HIP(hipGetLastError());
    const int ncolb = (nb + TN - 1) / TN;
    const long nrowb = (M + kNnRows - 1) / kNnRows;
    const long nwork = nrowb * ncolb;
    const long grid = (nwork + 7) / 8 * 8;
    hipLaunchKernelGGL((k_gemm_nn<BN>), dim3(grid), dim3(kNnThreads), 0, c.stream, A, Up, kpad, M,
                       nb, ncolb, D32, Out);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int split_gemm_tn(const float *A, const float *B, int nb, long K, double alpha, double *C,
                  int sym_result)
{
    return ctx().prec == 2 ? gemm_tn_split3(A, B, nb, K, alpha, C, sym_result != 0)
                           : gemm_tn_mode<float, 1>(A, B, nb, K, alpha, C, sym_result != 0);
}

template <class T>
int split_gemm_nn(const float *A, const double *U, int nb, long M, float *Out)
{
    return ctx().prec == 2 ? gemm_nn_split3(A, U, nb, M, Out) : gemm_nn_sgemm(A, U, nb, M, Out);
}

template int split_gemm_tn<float>(const float *, const float *, int, long, double, double *, int);
template int split_gemm_tn<double>(const float *, const float *, int, long, double, double *, int);
template int split_gemm_nn<float>(const float *, const double *, int, long, float *);
template int split_gemm_nn<double>(const float *, const double *, int, long, float *);

}  // namespace rsdft
