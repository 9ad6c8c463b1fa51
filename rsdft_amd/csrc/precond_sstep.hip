// s-step form of the preconditioner precond_hprsdft
// (cgpc_hprsdft_module.f90:126-263).
//
// The reference runs mloop CG steps on A x = g from x0 = g, where A is the
// 7-point -1/2 Laplacian of precond_cg_mat2_omp (E = 0):
//     r0 = g - A g, p0 = r0;  a_i = (r_i,r_i)/(p_i,A p_i);  x += a_i p_i;
//     r_{i+1} = r_i - a_i A p_i;  b_i = (r_{i+1},r_{i+1})/(r_i,r_i);  p_{i+1} = r_{i+1} + b_i p_i.
// Every r_i, p_i and x - g lies in the Krylov space of r0, so with a basis
//     v_0 = r0,  v_1 = tau(A) v_0,  v_{k+1} = 2 tau(A) v_k - v_{k-1},
//     tau(A) = (2/lmax) A - I,  lmax = 2 sum_d 1/h_d^2 >= spectrum of A
// (Chebyshev polynomials on the spectrum, so the basis stays well
// conditioned) the recurrence runs exactly in coefficient space: A acts as a
// tridiagonal matrix T on coefficients and every inner product is c^T G c'
// with G_jk = (v_j, v_k).  Pass 1 streams g once and accumulates the
// (m+1)x(m+1) Gram matrix per band; one thread per band then replays the
// reference's steps (same a_i, b_i, same early exit rr < 1e-24; it stops
// where r or p has vanished below the rounding of the Gram forms); pass 2
// streams g once more and writes Pg = g + sum_k d_k v_k.  Two reads and one
// write of psi-sized data replace ~30 passes of the step-by-step form.
//
// Both passes march z with the levels of the basis pipelined in registers:
// at step s level l (l = 0 is g, l = k+1 is v_k) has its newest plane at
// z = s - l; level l+1 at plane n uses level l at planes n-1, n, n+1 (z from
// registers, x and y from one LDS plane per level).  The y range shrinks by
// one row per level (redundant halo rows), the z range by one plane.
#include "common.h"

#include <algorithm>
#include <type_traits>

namespace rsdft {

__device__ __forceinline__ int wrapz(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

__device__ __forceinline__ long xcd_map(long L, long total) {
    if ((total & 7) != 0) return L;
    const long per = total >> 3;
    return (L & 7) * per + (L >> 3);
}

// Register rings of the levels.  Level l keeps its last rdepth(l) planes:
// three for the next level's stencil and back to plane s - NLEV where all
// levels are combined; level 0 (g) also holds the PF planes requested ahead.
// Every depth divides kRing and the z loop is unrolled kRing times, so plane
// z of level l sits in slot roff(l) + (z - s0) mod rdepth(l), a compile-time
// index in every unrolled step: no register moves between planes.
constexpr int kSstepPF = 2;
constexpr int kRing = 8;
// LDS row pitch (points): the row [pad, x = -1, x = 0 .. N1-1, x = N1, pad]
// takes N1 + 4; the pitch is rounded up to N1 modulo one LDS pass
__host__ __device__ constexpr int sstep_pitch(int N1, int elem)
{
    return N1 + (4 * elem + kLdsPassBytes - 1) / kLdsPassBytes * kLdsPassBytes / elem;
}
// Columns of the kernel's static reduction array (Gram entries in the Gram
// pass, one otherwise), and its bytes, which the dynamic LDS budget leaves out
__host__ __device__ constexpr int sstep_red_cols(int nlev, int mode)
{
    return mode == 0 ? nlev * (nlev + 1) / 2 + nlev : 1;
}
constexpr size_t sstep_static_lds(int nlev, int mode)
{
    return 16 * (size_t)sstep_red_cols(nlev, mode) * sizeof(double);
}
__host__ __device__ constexpr int rneed(int l, int NLEV) {
    return l == NLEV ? 1
                     : ((NLEV - l + 1) > 3 ? (NLEV - l + 1) : 3) + (l == 0 ? kSstepPF : 0);
}
__host__ __device__ constexpr int rdepth(int l, int NLEV) {
    return rneed(l, NLEV) == 1 ? 1 : (rneed(l, NLEV) <= 2 ? 2 : (rneed(l, NLEV) <= 4 ? 4 : 8));
}
__host__ __device__ constexpr int roff(int l, int NLEV) {
    return l == 0 ? 0 : roff(l - 1, NLEV) + rdepth(l - 1, NLEV);
}
// slot of level l for the plane d planes after s0 (d may be negative)
__host__ __device__ constexpr int rslot(int l, int NLEV, int d) {
    return roff(l, NLEV) + ((d % rdepth(l, NLEV)) + rdepth(l, NLEV)) % rdepth(l, NLEV);
}

// MODE 0: Gram pass, NLEV = m + 1, part[(e*nb+b)*tiles + tile] over the
//         (m+1)(m+2)/2 entries j <= k of the Gram matrix of v_0 .. v_m
//         (levels 1..NLEV, row-major upper triangle), then (g,g) and
//         (g, v_k) for k < m (level 0 against levels 0..NLEV-1).
// MODE 1: apply pass, NLEV = m, PG = g + sum_{k<m} d[b][k] v_k.
// MODE 2: apply pass fused with the CG direction update (cg_module.f90:258-
//         266): PG = g + sum_{k<m} d[b][k] v_k + beta[b] PG (beta == null:
//         PG is overwritten, the first CG step).
//
// Step s: level 0's current plane is s (requested kSstepPF steps earlier);
// level l = 1..NLEV computes plane s - l from level l-1 at planes s-l-1,
// s-l, s-l+1 (registers; s-l+1 was computed earlier in this step) with the
// x/y neighbours of plane s-l taken from LDS, where level l-1 left it at the
// end of step s-1.  Every level's newest plane is staged into one of two LDS
// buffers at the end of the step, so one barrier per plane suffices.
// workgroup size limit: fewer threads leave more registers per thread
// (FP64 768 threads: up to 168 registers; five levels need more)
template <class T, int NLEV>
constexpr int sstep_max_threads()
{
    return sizeof(T) == 8 ? (NLEV >= 5 ? 512 : 768) : (NLEV >= 5 ? 768 : 1024);
}
static int sstep_max_threads_rt(size_t elem, int nlev)
{
    return elem == 8 ? (nlev >= 5 ? 512 : 768) : (nlev >= 5 ? 768 : 1024);
}

template <class T, int PX, int NLEV, int MODE>
__global__ void __launch_bounds__((sstep_max_threads<T, NLEV>()))
k_pc_sstep(const T *__restrict__ G, T *__restrict__ PG, long ML, int N1, int N2, int N3,
           int ty, int tz, int tpr, double c0, double c1, double c2, double c3, double s1,
           int nb, const int *__restrict__ active, const double *__restrict__ dcoef,
           double *__restrict__ part, int tiles, const double *__restrict__ beta)
{
    constexpr int NG = sstep_red_cols(NLEV, 0);        // Gram entries
    constexpr int QTOT = roff(NLEV, NLEV) + rdepth(NLEV, NLEV);
    constexpr int PF = kSstepPF;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    T *Lbase = reinterpret_cast<T *>(smem_raw);
    __shared__ double red[16][sstep_red_cols(NLEV, MODE)];

    const long total = (long)tiles * nb;
    const long w = xcd_map(blockIdx.x, total);
    if (w >= total) return;
    const int band = (int)(w / tiles);
    const int tile = (int)(w - (long)band * tiles);
    if (!active[band]) return;
    const int nty = N2 / ty;
    const int y0 = (tile % nty) * ty;
    const int z0 = (tile / nty) * tz;
    const int rows = ty + 2 * NLEV;
    // LDS row: [pad, x = -1, x = 0 .. N1-1, x = N1, pad]; a thread's PX
    // adjacent points start at an even index, so they move as one vector
    const int W = sstep_pitch(N1, (int)sizeof(T));
    const int LS = rows * W;                 // one level plane in LDS
    const int tid = threadIdx.x;
    const int r = tid / tpr;                 // 0 .. rows-1
    const int x0 = (tid - r * tpr) * PX;
    const int gy = wrapz(y0 - NLEV + r, N2);
    const long plane = (long)N1 * N2;
    const T *src = G + (long)band * ML;
    const long rowoff = (long)gy * N1 + x0;
    const bool interior = r >= NLEV && r < NLEV + ty;
    typedef T VP __attribute__((ext_vector_type(PX)));

    // this thread's LDS positions (x0 - 1 of its row, and the rows above and
    // below) in every level plane of both buffers: each access below is an
    // immediate offset from one of them
    // (with five levels the row pointers are formed per step instead, to
    // stay within the register budget)
    constexpr bool kRowPtrs = NLEV <= 4;
    T *pc[2][NLEV], *pm[2][kRowPtrs ? NLEV : 1], *pp[2][kRowPtrs ? NLEV : 1];
#pragma unroll
    for (int bf = 0; bf < 2; ++bf)
#pragma unroll
        for (int l = 0; l < NLEV; ++l) {
            T *b0 = Lbase + (long)(bf * NLEV + l) * LS + r * W + 1 + x0;
            pc[bf][l] = b0;
            if constexpr (kRowPtrs) {
                pm[bf][l] = b0 - W;
                pp[bf][l] = b0 + W;
            }
        }

    double dk[NLEV > 0 ? NLEV : 1];
    if (MODE >= 1) {
#pragma unroll
        for (int k = 0; k < NLEV; ++k) dk[k] = dcoef[(long)band * NLEV + k];
    }
    const double bk = (MODE == 2 && beta) ? beta[band] : 0.0;
    // Gram partial sums of this thread's points, in the working precision.
    // In FP32 a thread adds at most 2 tz products per entry (a rounding error
    // ~ sqrt(2 tz) 2^-24 of its own partial, random in sign), and the ~10^4
    // partials of a band are summed in FP64 below, so an entry carries
    // ~1e-8 of |v_j||v_k|, below the rounding of the FP32 levels themselves;
    // FP32 accumulators take half the VALU cycles of FP64 ones
    T gacc[MODE == 0 ? NG : 1];
#pragma unroll
    for (int k = 0; k < (MODE == 0 ? NG : 1); ++k) gacc[k] = T(0);

    T q[QTOT][PX];
#pragma unroll
    for (int k = 0; k < QTOT; ++k)
#pragma unroll
        for (int px = 0; px < PX; ++px) q[k][px] = T(0);

    const int s0 = z0 - NLEV, s1e = z0 + tz + NLEV;
    // level 0: planes s0 - (depth-1-PF) .. s0 + PF - 1 (the prologue of the queue)
    {
        constexpr int D0 = rneed(0, NLEV);
#pragma unroll
        for (int k = 0; k < D0 - 1; ++k) {
            const int d = -(D0 - 1 - PF) + k;   // plane s0 + d
            const VP v = *reinterpret_cast<const VP *>(src + (long)wrapz(s0 + d, N3) * plane + rowoff);
#pragma unroll
            for (int px = 0; px < PX; ++px) q[rslot(0, NLEV, d)][px] = v[px];
        }
    }

    const T tc0 = (T)c0, tc1 = (T)c1, tc2 = (T)c2, tc3 = (T)c3, ts1 = (T)s1;
    // one plane step, s = s0 + U (mod kRing); CUR is the LDS buffer written at
    // its end ((s - s0) & 1), the other one holds what step s-1 left
    auto step = [&](auto u_tag, int s) {
        constexpr int U = decltype(u_tag)::value;
        constexpr int CUR = U & 1, PREV = CUR ^ 1;
        {   // request plane s + PF of g
            int zn = s + PF;
            while (zn < 0) zn += N3;
            while (zn >= N3) zn -= N3;
            const VP v = *reinterpret_cast<const VP *>(src + (long)zn * plane + rowoff);
#pragma unroll
            for (int px = 0; px < PX; ++px) q[rslot(0, NLEV, U + PF)][px] = v[px];
        }
#pragma unroll
        for (int l = 0; l < NLEV; ++l) {
            // level l+1 at plane s-l-1 on rows [l+1, rows-l-1): level l at planes
            // s-l-2 .. s-l; the neighbours across the chunk edges and rows of
            // plane s-l-1 from the previous buffer
            const int qc = rslot(l, NLEV, U - l - 1);
            const int qzm = rslot(l, NLEV, U - l - 2), qzp = rslot(l, NLEV, U - l);
            const int qn = rslot(l + 1, NLEV, U - l - 1);
            if (r >= l + 1 && r < rows - l - 1) {
                const T *c = pc[PREV][l];
                const T *cm, *cp;
                if constexpr (kRowPtrs) {
                    cm = pm[PREV][l];
                    cp = pp[PREV][l];
                } else {
                    cm = c - W;
                    cp = c + W;
                }
                const VP ym = *reinterpret_cast<const VP *>(cm + 1);
                const VP yp = *reinterpret_cast<const VP *>(cp + 1);
                const T xl = c[0], xr = c[PX + 1];
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    const T cen = q[qc][px];
                    const T left = px == 0 ? xl : q[qc][px - 1];
                    const T right = px == PX - 1 ? xr : q[qc][px + 1];
                    T a = tc0 * cen;
                    a = a + tc1 * (right + left) + tc2 * (yp[px] + ym[px]) +
                        tc3 * (q[qzp][px] + q[qzm][px]);
                    T v;
                    if (l == 0) {
                        v = cen - a;                          // v_0 = g - A g
                    } else if (l == 1) {
                        v = ts1 * a - cen;                    // v_1 = tau(A) v_0
                    } else {
                        // v_l = 2 tau(A) v_{l-1} - v_{l-2}; v_{l-2} is level l-1 at this plane
                        const T prev = q[rslot(l - 1, NLEV, U - l - 1)][px];
                        v = T(2) * (ts1 * a - cen) - prev;
                    }
                    q[qn][px] = v;
                }
            }
        }
        // all levels at plane zo = s - NLEV
        const int zo = s - NLEV;
        if (interior && zo >= z0) {
            if (MODE == 0) {
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    T vv[NLEV + 1];
#pragma unroll
                    for (int l = 0; l <= NLEV; ++l) vv[l] = q[rslot(l, NLEV, U - NLEV)][px];
                    int e = 0;
#pragma unroll
                    for (int j = 1; j <= NLEV; ++j)
#pragma unroll
                        for (int k = j; k <= NLEV; ++k)
                            gacc[e++] += vv[j] * vv[k];
#pragma unroll
                    for (int k = 0; k < NLEV; ++k)
                        gacc[e++] += vv[0] * vv[k];
                }
            } else {
                VP *dst = reinterpret_cast<VP *>(PG + (long)band * ML + (long)zo * plane + rowoff);
                VP old;
                if (MODE == 2 && beta) old = *dst;
                VP outv;
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    double x = (double)q[rslot(0, NLEV, U - NLEV)][px];
#pragma unroll
                    for (int k = 0; k < NLEV; ++k)
                        x += dk[k] * (double)q[rslot(k + 1, NLEV, U - NLEV)][px];
                    // MODE 2: P = Pg + beta P in the working precision, as k_cg_newp
                    T pv = (T)x;
                    if (MODE == 2 && beta) pv = pv + (T)bk * old[px];
                    outv[px] = pv;
                }
                *dst = outv;
            }
        }
        // stage the newest plane of levels 0 .. NLEV-1 for the next step
#pragma unroll
        for (int l = 0; l < NLEV; ++l) {
            if (r >= l && r < rows - l) {
                T *sc = pc[CUR][l] + 1;
                VP v;
#pragma unroll
                for (int px = 0; px < PX; ++px) v[px] = q[rslot(l, NLEV, U - l)][px];
                *reinterpret_cast<VP *>(sc) = v;
                if (x0 == 0) sc[N1] = v[0];                       // x = N1
                if (x0 + PX == N1) sc[-x0 - 1] = v[PX - 1];       // x = -1
            }
        }
        __syncthreads();
    };
    static_assert(kRing == 8, "the unrolled steps below");
    for (int s = s0; s < s1e; s += kRing) {
        step(std::integral_constant<int, 0>{}, s);
        if (s + 1 < s1e) step(std::integral_constant<int, 1>{}, s + 1);
        if (s + 2 < s1e) step(std::integral_constant<int, 2>{}, s + 2);
        if (s + 3 < s1e) step(std::integral_constant<int, 3>{}, s + 3);
        if (s + 4 < s1e) step(std::integral_constant<int, 4>{}, s + 4);
        if (s + 5 < s1e) step(std::integral_constant<int, 5>{}, s + 5);
        if (s + 6 < s1e) step(std::integral_constant<int, 6>{}, s + 6);
        if (s + 7 < s1e) step(std::integral_constant<int, 7>{}, s + 7);
    }
    if (MODE == 0) {
        const int lane = tid & 63, wave = tid >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int e = 0; e < NG; ++e) {
            double v = (double)gacc[e];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if (lane == 0) red[wave][e] = v;
        }
        __syncthreads();
        if (tid < NG) {
            double t = 0.0;
            for (int k = 0; k < nw; ++k) t += red[k][tid];
            part[((long)tid * nb + band) * tiles + tile] = t;
        }
    }
}

// Per band: Gram matrix -> the reference's mloop CG steps in coefficient
// space -> d[b][0..m-1] (Pg = g + sum d_k v_k); pc_active[b] = 0 when
// (r0,r0) < 1e-24 (the reference's early return), which gives d = 0.
// pgg[b] = (Pg, g) = (g,g) + sum_k d_k (v_k, g) from the same Gram matrix
// (the inner product the CG step takes next, cg_module.f90:237-243).
// Every inner product of the replay is a quadratic form c^T G c' whose
// rounding is about eps |c|^T |G| |c'| (eps: the precision the basis was
// computed in).  When the Krylov vectors of g are nearly dependent (a
// gradient concentrated on few eigenmodes of A: smooth, or a plane wave)
// the forms cancel to a small fraction `cond` of that bound and the replay
// loses log10(1/cond) digits; ill[b] = 1 then (cond < relmin), and the
// caller runs the step-by-step recurrence for that band instead.
template <int M>
__global__ void k_pc_sstep_scalars(const double *__restrict__ part, int nb, int tiles, double lmax,
                                   const int *__restrict__ active, double *__restrict__ dcoef,
                                   double *__restrict__ pgg, int *__restrict__ ill, double relmin)
{
    constexpr int N = M + 1;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !active[b]) return;
    double Gm[N][N], Gg[N];
    int e = 0;
    for (int j = 0; j < N; ++j)
        for (int k = j; k < N; ++k) {
            double t = 0.0;
            const double *p = part + ((long)e * nb + b) * tiles;
            for (int c = 0; c < tiles; ++c) t += p[c];
            Gm[j][k] = t;
            Gm[k][j] = t;
            ++e;
        }
    for (int k = 0; k < N; ++k) {   // Gg[0] = (g,g), Gg[k+1] = (g, v_k)
        double t = 0.0;
        const double *p = part + ((long)e * nb + b) * tiles;
        for (int c = 0; c < tiles; ++c) t += p[c];
        Gg[k] = t;
        ++e;
    }
    double d[M];
    for (int k = 0; k < M; ++k) d[k] = 0.0;
    double cond = 1.0;
    double rr1 = Gm[0][0];
    if (!(rr1 < 1.0e-24)) {
        double rc[N], pc[N], qc[N];
        for (int k = 0; k < N; ++k) rc[k] = pc[k] = 0.0;
        rc[0] = pc[0] = 1.0;
        const double h = 0.5 * lmax, qh = 0.25 * lmax;
        for (int il = 1; il <= M; ++il) {
            // q = T p  (A in the Chebyshev basis)
            for (int j = 0; j < N; ++j) qc[j] = 0.0;
            for (int k = 0; k < N; ++k) {
                if (pc[k] == 0.0) continue;
                qc[k] += h * pc[k];
                if (k == 0) {
                    if (N > 1) qc[1] += h * pc[0];
                } else {
                    qc[k - 1] += qh * pc[k];
                    if (k + 1 < N) qc[k + 1] += qh * pc[k];
                }
            }
            double pAp = 0.0, pAp_abs = 0.0;
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    const double t = pc[j] * Gm[j][k] * qc[k];
                    pAp += t;
                    pAp_abs += fabs(t);
                }
            cond = fmin(cond, fabs(pAp) / pAp_abs);
            const double rr0 = rr1;
            const double a = rr0 / pAp;
            for (int k = 0; k < M; ++k) d[k] += a * pc[k];
            if (il == M) break;
            for (int k = 0; k < N; ++k) rc[k] -= a * qc[k];
            rr1 = 0.0;
            double rr_abs = 0.0;
            for (int j = 0; j < N; ++j)
                for (int k = 0; k < N; ++k) {
                    const double t = rc[j] * Gm[j][k] * rc[k];
                    rr1 += t;
                    rr_abs += fabs(t);
                }
            cond = fmin(cond, fabs(rr1) / rr_abs);
            const double bb = rr1 / rr0;
            for (int k = 0; k < N; ++k) pc[k] = rc[k] + bb * pc[k];
        }
    }
    double pg = Gg[0];
    for (int k = 0; k < M; ++k) {
        dcoef[(long)b * M + k] = d[k];
        pg += d[k] * Gg[k + 1];
    }
    if (pgg) pgg[b] = pg;
    if (ill) ill[b] = cond < relmin ? 1 : 0;
}

struct SstepGeom {
    int ty, tz, tiles, threads, px;
    size_t lds;
};

// nlds: levels staged in LDS (two buffers each); the tallest tile that fits
// the thread and LDS limits (every level recomputes one halo row per side)
static SstepGeom sstep_geom(const GridDesc &g, int nlev, int nlds, int px, int maxthr,
                            size_t elem, size_t maxlds)
{
    SstepGeom G{};
    G.px = px;
    const int tpr = g.N1 / px;
    G.ty = 0;
    for (int ty = 1; ty <= g.N2; ++ty) {
        if (g.N2 % ty) continue;
        const int thr = tpr * (ty + 2 * nlev);
        if (thr > maxthr) break;
        const size_t lds = 2 * (size_t)nlds * (ty + 2 * nlev) * sstep_pitch(g.N1, (int)elem) * elem;
        if (lds > maxlds) break;
        G.ty = ty;
        G.threads = thr;
        G.lds = lds;
    }
    // the whole z column when short (the 2*nlev warm-up planes are paid once)
    int tz = g.N3;
    if (g.N3 > 96)
        for (int d = 96; d >= 1; --d)
            if (g.N3 % d == 0) { tz = d; break; }
    G.tz = tz;
    G.tiles = G.ty ? (g.N2 / G.ty) * (g.N3 / G.tz) : 0;
    return G;
}

template <class K>
static int allow_lds(K kernel, size_t lds)
{
    if (lds > 64 * 1024)
        RS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return kOK;
}

struct SstepConst {
    double c0, c1, c2, c3, lmax, s1;
};

static SstepConst sstep_const(const GridDesc &gd)
{
    SstepConst k;
    const double hx = 1.0 / (gd.h[0] * gd.h[0]), hy = 1.0 / (gd.h[1] * gd.h[1]),
                 hz = 1.0 / (gd.h[2] * gd.h[2]);
    k.c0 = hx + hy + hz;
    k.c1 = -0.5 * hx;
    k.c2 = -0.5 * hy;
    k.c3 = -0.5 * hz;
    k.lmax = 2.0 * k.c0;
    k.s1 = 2.0 / k.lmax;
    return k;
}

// pass 1: Gram of g, v_0 .. v_M, then the coefficients d and (Pg, g)
template <class T, int M>
static int sstep_coeffs(const T *g, int nb, const int *active, double *part, double *dcoef,
                        double *pgg, int *ill)
{
    Context &c = ctx();
    const GridDesc &gd = c.g;
    constexpr int PX = 2;
    const SstepConst k = sstep_const(gd);
    const SstepGeom G = sstep_geom(gd, M + 1, M + 1, PX, sstep_max_threads<T, M + 1>(),
                                   sizeof(T), c.max_lds - sstep_static_lds(M + 1, 0));
    if (G.ty == 0) return set_error(kErrArg, "no s-step tiling");
    const size_t lds = G.lds;
    const long nblk = (long)G.tiles * nb;
    RS_TRY(allow_lds(k_pc_sstep<T, PX, M + 1, 0>, lds));
    hipLaunchKernelGGL((k_pc_sstep<T, PX, M + 1, 0>), dim3((nblk + 7) / 8 * 8), dim3(G.threads),
                       lds, c.stream, g, (T *)nullptr, gd.ML, gd.N1, gd.N2, gd.N3, G.ty, G.tz,
                       gd.N1 / PX, k.c0, k.c1, k.c2, k.c3, k.s1, nb, active,
                       (const double *)nullptr, part, G.tiles, (const double *)nullptr);
    RS_HIP(hipGetLastError());
    // the replay keeps 8 digits of (Pg, g) and of Pg's coefficients in FP64; an
    // FP32 basis carries rounding ~6e-8 of each vector, so 1e-5 keeps 2 digits
    // of a direction that FP32 storage rounds anyway
    const double relmin = sizeof(T) == 8 ? 1.0e-8 : 1.0e-5;
    hipLaunchKernelGGL((k_pc_sstep_scalars<M>), dim3((nb + 127) / 128), dim3(128), 0, c.stream,
                       part, nb, G.tiles, k.lmax, active, dcoef, pgg, ill, relmin);
    RS_HIP(hipGetLastError());
    return kOK;
}

// pass 2: out = g + sum d_k v_k (fuse == 0), or out = that + beta out (fuse == 1)
template <class T, int M>
static int sstep_apply(const T *g, T *out, int nb, const int *active, const double *dcoef,
                       int fuse, const double *beta)
{
    Context &c = ctx();
    const GridDesc &gd = c.g;
    constexpr int PX = 2;
    const SstepConst k = sstep_const(gd);
    const SstepGeom G = sstep_geom(gd, M, M, PX, sstep_max_threads<T, M>(), sizeof(T),
                                   c.max_lds - sstep_static_lds(M, 1));
    if (G.ty == 0) return set_error(kErrArg, "no s-step tiling");
    const size_t lds = G.lds;
    const long nblk = (long)G.tiles * nb;
    RS_TRY(allow_lds(k_pc_sstep<T, PX, M, 1>, lds));
    RS_TRY(allow_lds(k_pc_sstep<T, PX, M, 2>, lds));
    if (fuse)
        hipLaunchKernelGGL((k_pc_sstep<T, PX, M, 2>), dim3((nblk + 7) / 8 * 8), dim3(G.threads),
                           lds, c.stream, g, out, gd.ML, gd.N1, gd.N2, gd.N3, G.ty, G.tz,
                           gd.N1 / PX, k.c0, k.c1, k.c2, k.c3, k.s1, nb, active, dcoef,
                           (double *)nullptr, G.tiles, beta);
    else
        hipLaunchKernelGGL((k_pc_sstep<T, PX, M, 1>), dim3((nblk + 7) / 8 * 8), dim3(G.threads),
                           lds, c.stream, g, out, gd.ML, gd.N1, gd.N2, gd.N3, G.ty, G.tz,
                           gd.N1 / PX, k.c0, k.c1, k.c2, k.c3, k.s1, nb, active, dcoef,
                           (double *)nullptr, G.tiles, (const double *)nullptr);
    RS_HIP(hipGetLastError());
    return kOK;
}

size_t precond_sstep_part_elems(int nb)
{
    Context &c = ctx();
    size_t mx = 0;
    for (int m = 1; m <= 4; ++m) {
        const size_t budget = c.max_lds - sstep_static_lds(m + 1, 0);
        const SstepGeom G = sstep_geom(c.g, m + 1, m + 1, 2, sstep_max_threads_rt(8, m + 1), 8,
                                       budget);
        const SstepGeom G4 = sstep_geom(c.g, m + 1, m + 1, 2, sstep_max_threads_rt(4, m + 1), 4,
                                        budget);
        const size_t t = (size_t)std::max(G.tiles, G4.tiles);
        mx = std::max(mx, t * (size_t)((m + 1) * (m + 2) / 2 + m + 1) * nb);
    }
    return mx + (size_t)4 * nb;
}

template <class T>
int precond_sstep_coeffs(const T *g, int nb, const int *active, double *part, double *dcoef,
                         double *pgg, int *ill)
{
    switch (ctx().mloop) {
    case 1: return sstep_coeffs<T, 1>(g, nb, active, part, dcoef, pgg, ill);
    case 2: return sstep_coeffs<T, 2>(g, nb, active, part, dcoef, pgg, ill);
    case 3: return sstep_coeffs<T, 3>(g, nb, active, part, dcoef, pgg, ill);
    case 4: return sstep_coeffs<T, 4>(g, nb, active, part, dcoef, pgg, ill);
    default: return set_error(kErrArg, "s-step preconditioner built for mloop 1..4");
    }
}

template <class T>
int precond_sstep_apply(const T *g, T *out, int nb, const int *active, const double *dcoef,
                        int fuse, const double *beta)
{
    switch (ctx().mloop) {
    case 1: return sstep_apply<T, 1>(g, out, nb, active, dcoef, fuse, beta);
    case 2: return sstep_apply<T, 2>(g, out, nb, active, dcoef, fuse, beta);
    case 3: return sstep_apply<T, 3>(g, out, nb, active, dcoef, fuse, beta);
    case 4: return sstep_apply<T, 4>(g, out, nb, active, dcoef, fuse, beta);
    default: return set_error(kErrArg, "s-step preconditioner built for mloop 1..4");
    }
}

template int precond_sstep_coeffs<double>(const double *, int, const int *, double *, double *,
                                          double *, int *);
template int precond_sstep_coeffs<float>(const float *, int, const int *, double *, double *,
                                         double *, int *);
template int precond_sstep_apply<double>(const double *, double *, int, const int *,
                                         const double *, int, const double *);
template int precond_sstep_apply<float>(const float *, float *, int, const int *, const double *,
                                        int, const double *);

}  // namespace rsdft
