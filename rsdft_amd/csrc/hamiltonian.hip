// Hamiltonian apply for gfx950: FD kinetic stencil + local potential +
// Kleinman-Bylander nonlocal projectors, and the 7-point preconditioner
// operator.
//
// Stencil design (one band per workgroup column, kernels in stencil_kernels.h):
//   * a workgroup owns `ty` full x-rows of the (x,y) plane and marches `tz`
//     planes in z; every thread owns 2 x points of one row (FP32: a
//     consecutive pair moved as one 8-byte access; FP64: two points N1/2
//     apart, so lanes touch consecutive words);
//   * the 2*Md+1 z-neighbours of a thread's points live in a register queue,
//     loaded PF planes ahead, so z costs no LDS traffic;
//   * the current plane (own rows + Md-deep y halo, x wrapped into pads) is
//     staged in one of two LDS buffers for the x and y terms: one barrier per
//     plane;
//   * periodic wrap in y/z is modular indexing of the global loads: the
//     reference's bcset_3 halo exchange (bcset_3_module.f90:15-151,
//     bc_module.f90:722-736) has no counterpart;
//   * the epilogue adds Vloc*psi (op_localpot);
//   * H psi: the nonlocal term is first written at every grid point by
//     k_nl_fill (grid order, coefficients of each range's projector groups
//     staged in LDS) and the stencil accumulates onto it, so the output is
//     streamed twice instead of read-modified-written in the partial cache
//     lines under the spheres; op_nonlocal (and layouts the fill does not
//     hold) adds the term with k_nl_apply, which visits only the grid points
//     inside the projector spheres, a group of projectors sharing a sphere at
//     a time, for many bands per table load.
#include "common.h"

#include <algorithm>

namespace rsdft {

}  // namespace rsdft

#include "stencil_kernels.h"

namespace rsdft {

template <class T, int P>
struct SlotVec {
    T v[P];
};

// ---------------------------------------------------------------------------
// Nonlocal projection (ps_nloc2_op_module.f90:161-175):
//   coefC[lma][b] = iuV(lma)*dV * sum_j uV(j,lma) psi_b(JJP(j,lma))
// One wave per (projector group, NBB bands): each grid value of the group's
// union point list is loaded once and multiplied into all P channels, FP64
// accumulation.  The 64 lanes' partial sums are added by a fixed butterfly,
// so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
// Sum of v[0..K) over the 64 lanes of a wave, K a power of two <= 64, by
// recursive halving: at distance d every lane keeps the half of its values
// that its side of the d-split owns and adds the partner's copy of that half,
// so after log2(K) steps lane l holds value l / (64/K) summed over 64/(64/K)
// lanes; the remaining distances complete the sum.  Fixed order, no LDS.
template <int D, int N, int K>
__device__ __forceinline__ void reduce_scatter_step(double (&v)[K], int lane)
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
        reduce_scatter_step<D / 2, H, K>(v, lane);
    }
}

template <int K>
__device__ __forceinline__ double wave_reduce_scatter(double (&v)[K])
{
    const int lane = threadIdx.x & 63;
    reduce_scatter_step<32, K, K>(v, lane);
    double r = v[0];
#pragma unroll
    for (int d = 32 / K; d > 0; d >>= 1) r += __shfl_xor(r, d, 64);
    return r;
}

template <class T, int P, int NBB>
__global__ void __launch_bounds__(64)
k_nl_project(const T *__restrict__ in, long ML, int nb, int ng, const int *__restrict__ gptr,
             const int *__restrict__ gidx, const T *__restrict__ gval, const int *__restrict__ glma,
             const double *__restrict__ scale, T *__restrict__ coefC,
             const int *__restrict__ band_active)
{
    constexpr int K = P * NBB;
    static_assert(K <= 64 && (K & (K - 1)) == 0, "P*NBB must be a power of two <= 64");
    const int nbblk = (nb + NBB - 1) / NBB;
    const long total = (long)ng * nbblk;
    const long w = xcd_work_index(blockIdx.x, total);
    if (w >= total) return;
    const int g = (int)(w % ng);
    const int b0 = (int)(w / ng) * NBB;
    const int lane = threadIdx.x;
    bool act[NBB];
#pragma unroll
    for (int k = 0; k < NBB; ++k) act[k] = (b0 + k < nb) && (!band_active || band_active[b0 + k]);
    double acc[K];
#pragma unroll
    for (int t = 0; t < K; ++t) acc[t] = 0.0;
    const int j1 = gptr[g + 1];
    for (int j = gptr[g] + lane; j < j1; j += 64) {
        const long i = gidx[j];
        SlotVec<T, P> u;
#pragma unroll
        for (int s = 0; s < P; ++s) u.v[s] = gval[(long)j * P + s];
#pragma unroll
        for (int k = 0; k < NBB; ++k) {
            if (!act[k]) continue;
            const double x = (double)in[(long)(b0 + k) * ML + i];
#pragma unroll
            for (int s = 0; s < P; ++s) acc[s * NBB + k] += (double)u.v[s] * x;
        }
    }
    // lane l ends with the total of value t = l / (64/K)
    const double sum = wave_reduce_scatter<K>(acc);
    constexpr int R = 64 / K;
    if (lane % R == 0) {
        const int t = lane / R;
        const int s = t / NBB, k = t - s * NBB;
        const int lma = glma[g * P + s];
        const int b = b0 + k;
        if (lma >= 0 && b < nb && (!band_active || band_active[b]))
            coefC[(long)lma * nb + b] = (T)(scale[lma] * sum);
    }
}

// ---------------------------------------------------------------------------
// Nonlocal apply (ps_nloc2_op_module.f90:255-263):
//   htpsi(i,b) += sum_{lma covering i} uV(i,lma) * coefC[lma][b]
// per group of one colour (disjoint point lists, so no two workgroups of a
// launch touch the same output element): every thread owns grid points of
// the group and adds the group's P channels for NBB bands.
// ---------------------------------------------------------------------------
template <class T, int P, int NBB>
__global__ void __launch_bounds__(256)
k_nl_apply(T *__restrict__ out, long ML, int nb, const int *__restrict__ order, int ngc,
           const int *__restrict__ gptr, const int *__restrict__ gidx, const T *__restrict__ gval,
           const int *__restrict__ glma, const T *__restrict__ coefC,
           const int *__restrict__ band_active)
{
    const int nbblk = (nb + NBB - 1) / NBB;
    const long total = (long)ngc * nbblk;
    const long w = xcd_work_index(blockIdx.x, total);
    if (w >= total) return;
    const int g = order[w % ngc];
    const int b0 = (int)(w / ngc) * NBB;
    bool act[NBB];
#pragma unroll
    for (int k = 0; k < NBB; ++k) act[k] = (b0 + k < nb) && (!band_active || band_active[b0 + k]);
    // the P x NBB coefficients are uniform; reading them back from LDS keeps
    // them in VGPRs (as scalars they would not fit the SGPR file)
    __shared__ T Lcf[P * NBB];
    for (int t = threadIdx.x; t < P * NBB; t += blockDim.x) {
        const int s = t / NBB, k = t - s * NBB;
        const int lma = glma[g * P + s];
        Lcf[t] = (lma >= 0 && b0 + k < nb && (!band_active || band_active[b0 + k]))
                     ? coefC[(long)lma * nb + b0 + k] : T(0);
    }
    __syncthreads();
    T cf[P][NBB];
#pragma unroll
    for (int s = 0; s < P; ++s)
#pragma unroll
        for (int k = 0; k < NBB; ++k) cf[s][k] = Lcf[s * NBB + k];
    const int j1 = gptr[g + 1];
    for (int j = gptr[g] + threadIdx.x; j < j1; j += blockDim.x) {
        const long i = gidx[j];
        SlotVec<T, P> u;
#pragma unroll
        for (int s = 0; s < P; ++s) u.v[s] = gval[(long)j * P + s];
#pragma unroll
        for (int k = 0; k < NBB; ++k) {
            if (!act[k]) continue;
            T a = T(0);
#pragma unroll
            for (int s = 0; s < P; ++s) a = a + u.v[s] * cf[s][k];
            T *o = out + (long)(b0 + k) * ML + i;
            *o = *o + a;
        }
    }
}

// ---------------------------------------------------------------------------
// Nonlocal term written for every grid point (H psi), in grid order:
//   out[b][i] = sum_{groups g covering i} sum_s uV(i, g, s) coefC[lma(g, s)][b]
// (0 outside every sphere); the stencil then accumulates kinetic + local onto
// it, so each output element is written by two streaming passes instead of a
// read-modify-write of the partial cache lines under the spheres.  One
// workgroup per (range of kFillRange points, NBK bands): the coefficients of
// the range's groups are staged in LDS ([slot][s][band], rows padded by 4
// elements so different slots start in different banks), every thread owns 4
// consecutive points (one 4-element vector store per band) and reads 4 bands
// per LDS access.
// ---------------------------------------------------------------------------
template <class T, int P, int MM, int NBK>
__global__ void __launch_bounds__(256)
k_nl_fill(T *__restrict__ out, long ML, int nb, int nrange, const int *__restrict__ rptr,
          const int *__restrict__ rgrp, const int *__restrict__ glma,
          const T *__restrict__ coefC, const unsigned *__restrict__ meta,
          const T *__restrict__ uv, const int *__restrict__ band_active)
{
    using V4 = T __attribute__((ext_vector_type(4)));
    constexpr int PT = 4, LS = P * NBK + 4;
    static_assert(kFillRange == 256 * PT, "one point quad per thread");
    extern __shared__ __attribute__((aligned(16))) unsigned char fill_raw[];
    T *Lc = reinterpret_cast<T *>(fill_raw);
    const int nbb = (nb + NBK - 1) / NBK;
    const long total = (long)nrange * nbb;
    const long w = xcd_work_index(blockIdx.x, total);
    if (w >= total) return;
    const int r = (int)(w % nrange);
    const int b0 = (int)(w / nrange) * NBK;
    if (band_active) {   // the same decision in every thread: no band of the block active
        bool any = false;
        for (int k = 0; k < NBK && b0 + k < nb; ++k) any = any || band_active[b0 + k] != 0;
        if (!any) return;
    }
    const int g0 = rptr[r], ngr = rptr[r + 1] - g0;
    for (int t = threadIdx.x; t < ngr * P * NBK; t += blockDim.x) {
        const int slot = t / (P * NBK), rem = t - slot * (P * NBK);
        const int s = rem / NBK, k = rem - s * NBK;
        const int lma = glma[rgrp[g0 + slot] * P + s];
        const int b = b0 + k;
        Lc[slot * LS + rem] = (lma >= 0 && b < nb) ? coefC[(long)lma * nb + b] : T(0);
    }
    __syncthreads();
    const long i0 = (long)r * kFillRange + threadIdx.x * PT;
    if (i0 >= ML) return;   // ML % 4 == 0: the quad is whole
    const uint4 mq = *reinterpret_cast<const uint4 *>(meta + i0);
    const unsigned mt[PT] = {mq.x, mq.y, mq.z, mq.w};
    int cnt[PT], cmax = 0;
#pragma unroll
    for (int p = 0; p < PT; ++p) {
        cnt[p] = (int)(mt[p] & 7u);
        cmax = max(cmax, cnt[p]);
    }
    T u[MM][PT * P];
#pragma unroll
    for (int m = 0; m < MM; ++m) {
        const V4 *src = reinterpret_cast<const V4 *>(uv + ((long)m * ML + i0) * P);
#pragma unroll
        for (int q = 0; q < PT * P / 4; ++q) {
            const V4 v = m < cmax ? src[q] : V4{T(0), T(0), T(0), T(0)};
            u[m][4 * q] = v.x;
            u[m][4 * q + 1] = v.y;
            u[m][4 * q + 2] = v.z;
            u[m][4 * q + 3] = v.w;
        }
    }
    for (int k0 = 0; k0 < NBK; k0 += 4) {
        T acc[PT][4];
#pragma unroll
        for (int p = 0; p < PT; ++p)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[p][j] = T(0);
#pragma unroll
        for (int p = 0; p < PT; ++p)
#pragma unroll
            for (int m = 0; m < MM; ++m) {
                if (m >= cnt[p]) continue;
                const int base = (int)((mt[p] >> (3 + 6 * m)) & 63u) * LS + k0;
#pragma unroll
                for (int s = 0; s < P; ++s) {
                    const V4 cv = *reinterpret_cast<const V4 *>(&Lc[base + s * NBK]);
                    const T uu = u[m][p * P + s];
                    acc[p][0] += uu * cv.x;
                    acc[p][1] += uu * cv.y;
                    acc[p][2] += uu * cv.z;
                    acc[p][3] += uu * cv.w;
                }
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int b = b0 + k0 + j;
            if (b >= nb || (band_active && !band_active[b])) continue;
            *reinterpret_cast<V4 *>(out + (long)b * ML + i0) =
                V4{acc[0][j], acc[1][j], acc[2][j], acc[3][j]};
        }
    }
}

// ---------------------------------------------------------------------------
// 7-point preconditioner operator of precond_cg_mat2_omp with E = 0
// (cgpc_hprsdft_module.f90:281-348):
//   (A f)(i) = c f(i) + sum_d c_d (f(i+e_d) + f(i-e_d)),
//   c = sum_d 1/h_d^2, c_d = -0.5/h_d^2.
// mode 0:  f = g:   Q = A g; R = g - Q; PP = R; PG = g; part = sum R^2
// mode 1:  f = PP:  Q = A PP;                    part = sum PP*Q
// Same lane-strided z-march as k_hpsi with a 1-plane prefetch.
// ---------------------------------------------------------------------------
template <class T, int PX>
__global__ void __launch_bounds__(512)
k_pc_stencil(const T *__restrict__ in, T *__restrict__ Q, T *__restrict__ R,
             T *__restrict__ PP, T *__restrict__ PG, long ML, int N1, int N2, int N3, int ty,
             int tz, int tpr, T c0, T c1, T c2, T c3, int mode,
             const int *__restrict__ band_active, double *__restrict__ part, int tiles, int nbands)
{
    extern __shared__ unsigned char smem_raw[];
    T *L = reinterpret_cast<T *>(smem_raw);
    __shared__ double red[8];
    const long total = (long)tiles * nbands;
    const long wk = xcd_work_index(blockIdx.x, total);
    if (wk >= total) return;
    const int band = (int)(wk / tiles);
    const int tile = (int)(wk - (long)band * tiles);
    if (band_active && !band_active[band]) return;
    const int nty = N2 / ty;
    const int y0 = (tile % nty) * ty;
    const int z0 = (tile / nty) * tz;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int r = tid / tpr;
    const int cx0 = tid - r * tpr;
    const int yrow = y0 + r;
    const long plane = (long)N1 * N2;
    const long boff = (long)band * ML;
    const T *src = in + boff;
    const long rowoff = (long)yrow * N1 + cx0;
    // halo: rows y0-1 and y0+ty, 2*tpr items, at most one per thread (ty >= 2 or tpr*ty >= 2*tpr)
    const int nhalo = 2 * tpr;
    const bool h_on = tid < nhalo;
    const int hh = h_on ? tid / tpr : 0;
    const int hc = tid - hh * tpr;
    const int h_lrow = (hh == 0 ? 0 : ty + 1) * N1 + hc;
    const int h_goff = wrapi(hh == 0 ? y0 - 1 : y0 + ty, N2) * N1 + hc;

    T qm[PX], q0[PX], qp[PX], qn[PX], hb[PX];
    {
        const T *pm = src + wrapi(z0 - 1, N3) * plane + rowoff;
        const T *p0 = src + (long)z0 * plane + rowoff;
        const T *pp = src + wrapi(z0 + 1, N3) * plane + rowoff;
#pragma unroll
        for (int px = 0; px < PX; ++px) {
            qm[px] = pm[px * tpr];
            q0[px] = p0[px * tpr];
            qp[px] = pp[px * tpr];
        }
        if (h_on) {
            const T *ph = src + (long)z0 * plane + h_goff;
#pragma unroll
            for (int px = 0; px < PX; ++px) hb[px] = ph[px * tpr];
        }
    }
    T *Lrow = L + (r + 1) * N1 + cx0;
    double sum = 0.0;
    for (int z = z0; z < z0 + tz; ++z) {
        {
            int zn = z + 2;
            if (zn >= N3) zn -= N3;
            const T *p = src + zn * plane + rowoff;
#pragma unroll
            for (int px = 0; px < PX; ++px) qn[px] = p[px * tpr];
        }
#pragma unroll
        for (int px = 0; px < PX; ++px) Lrow[px * tpr] = q0[px];
        if (h_on) {
#pragma unroll
            for (int px = 0; px < PX; ++px) L[h_lrow + px * tpr] = hb[px];
        }
        __syncthreads();
        if (h_on) {
            int zn = z + 1;
            if (zn >= N3) zn -= N3;
            const T *p = src + zn * plane + h_goff;
#pragma unroll
            for (int px = 0; px < PX; ++px) hb[px] = p[px * tpr];
        }
#pragma unroll
        for (int px = 0; px < PX; ++px) {
            const int x = cx0 + px * tpr;
            const int dm = x == 0 ? N1 - 1 : -1;
            const int dp = x == N1 - 1 ? 1 - N1 : 1;
            const T *lr = Lrow + px * tpr;
            T f = c0 * q0[px];
            f = f + c1 * (lr[dp] + lr[dm]) + c2 * (lr[N1] + lr[-N1]) + c3 * (qp[px] + qm[px]);
            const long gi = boff + (long)z * plane + rowoff + px * tpr;
            if (mode == 0) {
                const T rr = q0[px] - f;
                R[gi] = rr;
                PP[gi] = rr;
                PG[gi] = q0[px];
                sum += (double)rr * (double)rr;
            } else {
                Q[gi] = f;
                sum += (double)q0[px] * (double)f;
            }
        }
        __syncthreads();
#pragma unroll
        for (int px = 0; px < PX; ++px) {
            qm[px] = q0[px];
            q0[px] = qp[px];
            qp[px] = qn[px];
        }
    }
    const int lane = tid & 63, wave = tid >> 6, nw = (nthr + 63) >> 6;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
    if (lane == 0) red[wave] = sum;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[w];
        part[(long)band * tiles + tile] = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
StencilGeom choose_geom(int N1, int N2, int N3, int md, int px, int elem_bytes)
{
    StencilGeom g;
    g.px = px;
    const int tpr = N1 / px;
    double best = -1.0;
    g.threads = 0;
    for (int ty = md; ty <= N2; ++ty) {   // ty >= md keeps halo items <= 2 per thread
        if (N2 % ty) continue;
        const int thr = tpr * ty;
        if (thr > 512) break;
        const size_t lds = (size_t)(ty + 2 * md) * (N1 + 2 * md) * elem_bytes;
        if (lds > 48 * 1024) continue;
        const int waves = (thr + 63) / 64;
        const double eff = (double)thr / (64.0 * waves);
        // favour fuller waves, then taller tiles (less halo), then >= 128 threads
        const double score = eff * 4.0 + (double)ty / (ty + 2.0 * md) + (thr >= 128 ? 1.0 : 0.0);
        if (score > best) {
            best = score;
            g.ty = ty;
            g.threads = thr;
        }
    }
    // z-march length: long enough to amortise the 2*Md queue warm-up
    int tz = N3;
    for (int c = 1; c <= N3; ++c) {
        if (N3 % c) continue;
        if (N3 / c >= 4 * md + 8 || c == 1) tz = N3 / c;
        if (N3 / c < 4 * md + 8) break;
    }
    if (tz > 48) {
        for (int c = 2; c <= N3; ++c)
            if (N3 % c == 0 && N3 / c <= 48 && N3 / c >= 2 * md + 4) { tz = N3 / c; break; }
    }
    g.tz = tz;
    g.tiles = (N2 / g.ty) * (N3 / g.tz);
    return g;
}

StencilGeom choose_hpsi_geom(int N1, int N2, int N3, int md, int px, int pad, int elem_bytes,
                             int max_threads)
{
    StencilGeom g;
    g.px = px;
    g.threads = 0;
    if (N1 % px) return g;
    const int tpr = N1 / px;
    double best = -1.0;
    for (int ty = md; ty <= N2; ++ty) {   // ty >= md keeps halo items <= 2 per thread
        if (N2 % ty) continue;
        const int thr = tpr * ty;
        if (thr > max_threads) break;
        // LDS rows: N1 + 2*pad points for the strided FP64 kernel, the padded
        // chunk pitch for the chunked FP32 kernel
        const int row = elem_bytes == 8 ? N1 + 2 * pad
                                        : hpsi2_pitch(tpr, pad / px, px * elem_bytes) * px;
        const size_t lds = 2 * (size_t)(ty + 2 * md) * row * elem_bytes;
        if (lds > 64 * 1024) continue;
        const int waves = (thr + 63) / 64;
        const double eff = (double)thr / (64.0 * waves);
        // one halo item per thread first, then full waves, then taller tiles
        const double score = (ty >= 2 * md ? 8.0 : 0.0) + eff * 4.0 + (double)ty / (ty + 2.0 * md) +
                             (thr >= 128 ? 1.0 : 0.0);
        if (score > best) {
            best = score;
            g.ty = ty;
            g.threads = thr;
            g.lds = lds;
        }
    }
    g.hm = g.ty >= 2 * md ? 1 : 2;
    // z-march: the whole column when it is short, else the longest divisor <= 96
    int tz = N3;
    if (N3 > 96)
        for (int d = 96; d >= 1; --d)
            if (N3 % d == 0) { tz = d; break; }
    g.tz = tz;
    g.tiles = (N2 / g.ty) * (N3 / g.tz);
    return g;
}

template <class T, int MD, int HM, int FL>
static void launch_hpsi_k(const T *in, T *out, int nb, const int *band_active, const T *vloc,
                          const StencilCoef<T, MD> &cf)
{
    Context &c = ctx();
    const GridDesc &g = c.g;
    const StencilGeom &G = sizeof(T) == 8 ? c.geom_h : c.geom_h32;
    const long nblk = (long)G.tiles * nb;
    const dim3 grid((unsigned)((nblk + 7) / 8 * 8));
    if constexpr (sizeof(T) == 8)
        hipLaunchKernelGGL((k_hpsi_s<T, MD, kHpsiPx64, HM, FL, 1>), grid, dim3(G.threads), G.lds,
                           c.stream, in, out, g.ML, g.N1, g.N2, g.N3, G.ty, G.tz, G.tiles, nb, cf,
                           vloc, band_active);
    else
        hipLaunchKernelGGL((k_hpsi2<T, MD, kHpsiPx32, HM, FL, kHpsiThreads32>), grid,
                           dim3(G.threads), G.lds, c.stream, in, out, g.ML, g.N1, g.N2, g.N3, G.ty,
                           G.tz, G.tiles, nb, cf, vloc, band_active);
}

template <class T, int MD>
static int launch_hpsi_md(const T *in, T *out, int nb, int flags, const int *band_active,
                          const T *vloc)
{
    Context &c = ctx();
    const GridDesc &g = c.g;
    const StencilGeom &G = sizeof(T) == 8 ? c.geom_h : c.geom_h32;
    StencilCoef<T, MD> cf;
    cf.c0 = (T)g.coef_lap0;
    for (int m = 0; m < MD; ++m) {
        cf.cx[m] = (T)g.coef_lap[0][m];
        cf.cy[m] = (T)g.coef_lap[1][m];
        cf.cz[m] = (T)g.coef_lap[2][m];
    }
    const bool h1 = G.hm == 1;
    if (flags == (kKin | kLoc)) {
        if (h1) launch_hpsi_k<T, MD, 1, kKin | kLoc>(in, out, nb, band_active, vloc, cf);
        else launch_hpsi_k<T, MD, 2, kKin | kLoc>(in, out, nb, band_active, vloc, cf);
    } else if (flags == kKin) {
        if (h1) launch_hpsi_k<T, MD, 1, kKin>(in, out, nb, band_active, vloc, cf);
        else launch_hpsi_k<T, MD, 2, kKin>(in, out, nb, band_active, vloc, cf);
    } else if (flags == (kKin | kLoc | kAcc)) {
        if (h1) launch_hpsi_k<T, MD, 1, kKin | kLoc | kAcc>(in, out, nb, band_active, vloc, cf);
        else launch_hpsi_k<T, MD, 2, kKin | kLoc | kAcc>(in, out, nb, band_active, vloc, cf);
    } else {
        return set_error(kErrArg, "stencil flags %d not supported", flags);
    }
    RS_HIP(hipGetLastError());
    return kOK;
}

// out[b][i] += V[i] * in[b][i]  (op_localpot alone, localpot_module.f90:27-45)
template <class T>
__global__ void __launch_bounds__(256)
k_localpot_acc(const T *__restrict__ in, T *__restrict__ out, const T *__restrict__ vloc, long ML,
               int nb, const int *__restrict__ band_active)
{
    const long n = ML * nb;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const long b = i / ML;
        if (band_active && !band_active[b]) continue;
        out[i] = out[i] + vloc[i - b * ML] * in[i];
    }
}

template <class T>
static int launch_stencil(const T *in, T *out, int nb, int flags, const int *band_active)
{
    Context &c = ctx();
    const T *vloc = sizeof(T) == 8 ? (const T *)c.vloc64.p : (const T *)c.vloc32.p;
    if ((flags & kLoc) && !vloc) return set_error(kErrState, "local potential not set");
    if (flags == (kLoc | kAcc)) {
        const long n = c.g.ML * nb;
        const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
        hipLaunchKernelGGL((k_localpot_acc<T>), dim3(blocks), dim3(256), 0, c.stream, in, out,
                           vloc, c.g.ML, nb, band_active);
        RS_HIP(hipGetLastError());
        return kOK;
    }
    switch (c.g.Md) {
    case 1: return launch_hpsi_md<T, 1>(in, out, nb, flags, band_active, vloc);
    case 2: return launch_hpsi_md<T, 2>(in, out, nb, flags, band_active, vloc);
    case 3: return launch_hpsi_md<T, 3>(in, out, nb, flags, band_active, vloc);
    case 4: return launch_hpsi_md<T, 4>(in, out, nb, flags, band_active, vloc);
    case 6: return launch_hpsi_md<T, 6>(in, out, nb, flags, band_active, vloc);
    case 8: return launch_hpsi_md<T, 8>(in, out, nb, flags, band_active, vloc);
    default: return set_error(kErrArg, "Md=%d not compiled (1,2,3,4,6,8)", c.g.Md);
    }
}

template <class T, int P>
static void launch_nl_project(const T *in, int nb, T *coef, const int *band_active)
{
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    constexpr int NBB = P >= 16 ? 2 : 32 / P > 16 ? 16 : 32 / P;
    const long nblk = (long)nl.ng * ((nb + NBB - 1) / NBB);
    const T *val = sizeof(T) == 8 ? (const T *)nl.g_val64.p : (const T *)nl.g_val32.p;
    hipLaunchKernelGGL((k_nl_project<T, P, NBB>), dim3((unsigned)((nblk + 7) / 8 * 8)), dim3(64), 0,
                       c.stream, in, c.g.ML, nb, nl.ng, nl.g_ptr.as<int>(), nl.g_idx.as<int>(), val,
                       nl.g_lma.as<int>(), nl.scale64.as<double>(), coef, band_active);
}

template <class T, int P>
static void launch_nl_apply(T *out, int nb, const T *coef, const int *band_active)
{
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    // P*NBB coefficients per thread: 64 FP32 / 32 FP64 registers
    constexpr int NBB = sizeof(T) == 8 ? (P >= 8 ? 2 : 32 / P / 2 * 2 > 8 ? 8 : 32 / P)
                                       : (P >= 16 ? 4 : 64 / P > 16 ? 16 : 64 / P);
    const T *val = sizeof(T) == 8 ? (const T *)nl.g_val64.p : (const T *)nl.g_val32.p;
    for (size_t col = 0; col + 1 < nl.color_ptr.size(); ++col) {
        const int g0 = nl.color_ptr[col], ngc = nl.color_ptr[col + 1] - g0;
        if (ngc <= 0) continue;
        const long nblk = (long)ngc * ((nb + NBB - 1) / NBB);
        hipLaunchKernelGGL((k_nl_apply<T, P, NBB>), dim3((unsigned)((nblk + 7) / 8 * 8)), dim3(256),
                           0, c.stream, out, c.g.ML, nb, nl.g_order.as<int>() + g0, ngc,
                           nl.g_ptr.as<int>(), nl.g_idx.as<int>(), val, nl.g_lma.as<int>(), coef,
                           band_active);
    }
}

template <class T>
int nl_project(const T *in, int nb, T *coef, const int *band_active)
{
    Context &c = ctx();
    if (c.nl.nz <= 0) return kOK;
    switch (c.nl.gp) {
    case 1: launch_nl_project<T, 1>(in, nb, coef, band_active); break;
    case 2: launch_nl_project<T, 2>(in, nb, coef, band_active); break;
    case 4: launch_nl_project<T, 4>(in, nb, coef, band_active); break;
    case 8: launch_nl_project<T, 8>(in, nb, coef, band_active); break;
    case 16: launch_nl_project<T, 16>(in, nb, coef, band_active); break;
    default: return set_error(kErrState, "projector group width %d", c.nl.gp);
    }
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int nl_apply(T *out, int nb, const T *coef, const int *band_active)
{
    Context &c = ctx();
    if (c.nl.nz <= 0 || c.nl.npts <= 0) return kOK;
    switch (c.nl.gp) {
    case 1: launch_nl_apply<T, 1>(out, nb, coef, band_active); break;
    case 2: launch_nl_apply<T, 2>(out, nb, coef, band_active); break;
    case 4: launch_nl_apply<T, 4>(out, nb, coef, band_active); break;
    case 8: launch_nl_apply<T, 8>(out, nb, coef, band_active); break;
    case 16: launch_nl_apply<T, 16>(out, nb, coef, band_active); break;
    default: return set_error(kErrState, "projector group width %d", c.nl.gp);
    }
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T, int P, int MM>
static void launch_nl_fill(int nb, const T *coef, const int *band_active, T *out)
{
    Context &c = ctx();
    const NonlocalTables &nl = c.nl;
    // LDS: fill_gmax (<= 64) slots x (P NBK + 4) elements stays below 64 KB
    constexpr int NBK = sizeof(T) == 4 ? (P <= 4 ? 32 : 128 / P) : (P <= 4 ? 16 : 64 / P);
    constexpr int LS = P * NBK + 4;
    const long nblk = (long)nl.nrange * ((nb + NBK - 1) / NBK);
    const size_t lds = sizeof(T) * (size_t)std::max(nl.fill_gmax, 1) * LS;
    const T *uv = sizeof(T) == 8 ? (const T *)nl.fill_uv64.p : (const T *)nl.fill_uv.p;
    hipLaunchKernelGGL((k_nl_fill<T, P, MM, NBK>), dim3((unsigned)((nblk + 7) / 8 * 8)),
                       dim3(256), lds, c.stream, out, c.g.ML, nb, nl.nrange,
                       nl.fill_rptr.as<int>(), nl.fill_rgrp.as<int>(), nl.g_lma.as<int>(), coef,
                       nl.fill_meta.as<unsigned>(), uv, band_active);
}

// out[b][i] <- nonlocal term at every grid point of the active bands
template <class T>
static int nl_fill(int nb, const T *coef, const int *band_active, T *out)
{
    const NonlocalTables &nl = ctx().nl;
    switch (nl.gp * 8 + nl.fill_mm) {
#define RS_FILL_CASE(P, MM) \
    case P * 8 + MM: launch_nl_fill<T, P, MM>(nb, coef, band_active, out); break;
    RS_FILL_CASE(1, 1) RS_FILL_CASE(1, 2) RS_FILL_CASE(1, 4)
    RS_FILL_CASE(2, 1) RS_FILL_CASE(2, 2) RS_FILL_CASE(2, 4)
    RS_FILL_CASE(4, 1) RS_FILL_CASE(4, 2) RS_FILL_CASE(4, 4)
    RS_FILL_CASE(8, 1) RS_FILL_CASE(8, 2) RS_FILL_CASE(16, 1)
#undef RS_FILL_CASE
    default: return set_error(kErrState, "fill layout gp=%d mm=%d", nl.gp, nl.fill_mm);
    }
    RS_HIP(hipGetLastError());
    return kOK;
}

static bool fill_ready() { return ctx().nl.nz > 0 && ctx().nl.fill_mm > 0; }

// coef: room for nb * nz projection coefficients supplied by the caller, or
// null to use (and grow) the context's buffer
template <class T>
static int hpsi_apply_untimed(const T *in, T *out, int nb, int flags, const int *band_active,
                              T *coef)
{
    Context &c = ctx();
    if (c.nl.nz <= 0) flags &= ~kNl;
    const int sflags = flags & (kKin | kLoc | kAcc);
    if ((flags & kNl) && !coef) {
        RS_TRY(c.coefC.ensure((size_t)nb * c.nl.nz * sizeof(T)));
        coef = c.coefC.as<T>();
    }
    if (flags & kNl) RS_TRY(nl_project<T>(in, nb, coef, band_active));
    // H psi: the nonlocal term written for every grid point, then the stencil
    // accumulates kinetic + local onto it
    if (sflags == (kKin | kLoc) && (flags & kNl) && fill_ready()) {
        RS_TRY(nl_fill<T>(nb, coef, band_active, out));
        return launch_stencil<T>(in, out, nb, kKin | kLoc | kAcc, band_active);
    }
    if (sflags & (kKin | kLoc)) {
        RS_TRY(launch_stencil<T>(in, out, nb, sflags, band_active));
    } else if (!(flags & kAcc)) {
        RS_HIP(hipMemsetAsync(out, 0, sizeof(T) * (size_t)c.g.ML * nb, c.stream));
    }
    if (flags & kNl) RS_TRY(nl_apply<T>(out, nb, coef, band_active));
    return kOK;
}

template <class T>
int hpsi_apply(const T *in, T *out, int nb, int flags, const int *band_active, T *coef)
{
    if (nb <= 0) return kOK;
    int h;
    RS_TRY(phase_begin(kTHam, &h));
    RS_TRY(hpsi_apply_untimed<T>(in, out, nb, flags, band_active, coef));
    RS_TRY(phase_end(h));
    // bands applied: all nb, or those the mask selects (counted on the device)
    if (band_active) RS_TRY(count_active_bands(band_active, nb));
    else ctx().hpsi_bands += nb;
    return kOK;
}

template <class T>
int hpsi_stencil_only(const T *in, T *out, int nb)
{
    if (fill_ready()) return launch_stencil<T>(in, out, nb, kKin | kLoc | kAcc, nullptr);
    return launch_stencil<T>(in, out, nb, kKin | kLoc, nullptr);
}

template <class T>
int stencil_kinloc(const T *in, T *out, int nb)
{
    return launch_stencil<T>(in, out, nb, kKin | kLoc, nullptr);
}

template <class T>
int nl_stage(T *out, int nb, const T *coef)
{
    if (fill_ready()) return nl_fill<T>(nb, coef, nullptr, out);
    return nl_apply<T>(out, nb, coef, nullptr);
}

// Plain copy of nb bands with 16-byte vector loads and stores: the stream
// rate a kernel that reads one array and writes one attains on this device.
template <class T>
__global__ void __launch_bounds__(256) k_copy_bands(const T *__restrict__ in, T *__restrict__ out, long n)
{
    using V = uint4;   // 16 bytes
    constexpr int E = 16 / sizeof(T);
    const long nv = n / E;
    const V *vi = reinterpret_cast<const V *>(in);
    V *vo = reinterpret_cast<V *>(out);
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nv; i += stride) vo[i] = vi[i];
    for (long i = nv * E + blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = in[i];
}

template <class T>
int copy_bands(const T *in, T *out, int nb)
{
    Context &c = ctx();
    const long n = c.g.ML * nb;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15)
        return set_error(kErrArg, "copy_bands: 16-byte aligned arrays expected");
    hipLaunchKernelGGL((k_copy_bands<T>), dim3(8192), dim3(256), 0, c.stream, in, out, n);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int pc_stencil(const T *in, T *Q, T *R, T *PP, T *PG, int nb, int mode, const int *active,
               double *part)
{
    Context &c = ctx();
    const GridDesc &g = c.g;
    const StencilGeom &G = c.geom_pc;
    const double hx = 1.0 / (g.h[0] * g.h[0]), hy = 1.0 / (g.h[1] * g.h[1]),
                 hz = 1.0 / (g.h[2] * g.h[2]);
    const T c0 = (T)(hx + hy + hz), c1 = (T)(-0.5 * hx), c2 = (T)(-0.5 * hy), c3 = (T)(-0.5 * hz);
    const int tpr = g.N1 / G.px;
    const size_t lds = (size_t)(G.ty + 2) * g.N1 * sizeof(T);
    const long nblk = (long)G.tiles * nb;
    hipLaunchKernelGGL((k_pc_stencil<T, 4>), dim3((nblk + 7) / 8 * 8), dim3(G.threads), lds,
                       c.stream, in, Q, R, PP, PG, g.ML, g.N1, g.N2, g.N3, G.ty, G.tz, tpr, c0, c1,
                       c2, c3, mode, active, part, G.tiles, nb);
    RS_HIP(hipGetLastError());
    return kOK;
}

template int hpsi_apply<double>(const double *, double *, int, int, const int *, double *);
template int hpsi_apply<float>(const float *, float *, int, int, const int *, float *);
template int hpsi_stencil_only<double>(const double *, double *, int);
template int stencil_kinloc<double>(const double *, double *, int);
template int hpsi_stencil_only<float>(const float *, float *, int);
template int nl_stage<double>(double *, int, const double *);
template int nl_stage<float>(float *, int, const float *);
template int copy_bands<double>(const double *, double *, int);
template int copy_bands<float>(const float *, float *, int);
template int nl_project<double>(const double *, int, double *, const int *);
template int nl_apply<double>(double *, int, const double *, const int *);
template int nl_apply<float>(float *, int, const float *, const int *);
template int nl_project<float>(const float *, int, float *, const int *);
template int pc_stencil<double>(const double *, double *, double *, double *, double *, int,
                                int, const int *, double *);
template int pc_stencil<float>(const float *, float *, float *, float *, float *, int, int,
                               const int *, double *);

}  // namespace rsdft
