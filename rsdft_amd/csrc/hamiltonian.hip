// Hamiltonian apply for gfx950: FD kinetic stencil + local potential +
// Kleinman-Bylander nonlocal projectors, and the 7-point preconditioner
// operator.
//
// Stencil design (one band per workgroup column):
//   * a workgroup owns `ty` full x-rows of the (x,y) plane and marches `tz`
//     planes in z; every thread owns `PX` consecutive x points of one row;
//   * the 2*Md+1 z-neighbours of a thread's points live in a register queue
//     (one new plane loaded per step), so z costs no LDS traffic;
//   * the current plane (own rows + Md-deep y halo, full x so the x wrap is
//     a modular index into LDS) is staged in LDS for the x and y terms;
//   * periodic wrap in y/z is modular indexing of the global loads: the
//     reference's bcset_3 halo exchange (bcset_3_module.f90:15-151,
//     bc_module.f90:722-736) has no counterpart;
//   * the epilogue adds Vloc*psi (op_localpot);
//   * the nonlocal term is added afterwards by k_nl_apply, which visits only
//     the grid points inside some projector sphere, for many bands per
//     entry load.
#include "common.h"

#include <algorithm>

namespace rsdft {

}  // namespace rsdft

#include "stencil_kernels.h"

namespace rsdft {

// ---------------------------------------------------------------------------
// Nonlocal projection: coefC[lma][b] = iuV(lma)*dV*sum_j uV(j,lma) psi_b(JJP(j,lma))
// (ps_nloc2_op_module.f90:161-175), FP64 accumulation, NBB bands per block.
// ---------------------------------------------------------------------------
template <int NBB>
__device__ __forceinline__ void block_reduce_store(double (&acc)[NBB], double *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NBB; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[wave * NBB + k] = v;
    }
    __syncthreads();
    if (threadIdx.x < NBB) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += red[w * NBB + threadIdx.x];
        acc[0] = s;   // thread k holds the total of band k in acc[0]
    }
}

template <class T, int NBB>
__global__ void __launch_bounds__(256)
k_project(const T *__restrict__ in, long ML, int nb, int nz, const int *__restrict__ rptr,
          const int *__restrict__ ridx, const T *__restrict__ rval,
          const double *__restrict__ scale, T *__restrict__ coefC,
          const int *__restrict__ band_active)
{
    __shared__ double red[16 * NBB];
    const int nbblk = (nb + NBB - 1) / NBB;
    const long total = (long)nz * nbblk;
    const long w = xcd_work_index(blockIdx.x, total);
    const int lma = (int)(w % nz);
    const int b0 = (int)(w / nz) * NBB;
    double acc[NBB];
#pragma unroll
    for (int k = 0; k < NBB; ++k) acc[k] = 0.0;
    bool act[NBB];
#pragma unroll
    for (int k = 0; k < NBB; ++k)
        act[k] = (b0 + k < nb) && (!band_active || band_active[b0 + k]);
    const int j1 = rptr[lma + 1];
    for (int j = rptr[lma] + threadIdx.x; j < j1; j += blockDim.x) {
        const int i = ridx[j];
        const double v = (double)rval[j];
#pragma unroll
        for (int k = 0; k < NBB; ++k)
            if (act[k]) acc[k] += v * (double)in[(long)(b0 + k) * ML + i];
    }
    block_reduce_store<NBB>(acc, red);
    if (threadIdx.x < NBB) {
        const int b = b0 + threadIdx.x;
        if (b < nb && (!band_active || band_active[b]))
            coefC[(long)lma * nb + b] = (T)(scale[lma] * acc[0]);
    }
}

// ---------------------------------------------------------------------------
// Nonlocal apply (ps_nloc2_op_module.f90:255-263):
//   htpsi(i,b) += sum_{lma covering i} uV(i,lma) * coefC[lma][b]
// over the grid points that lie in some sphere ("covered" points).  A
// workgroup stages the point-major entries (lma, uV) of its 256 points in LDS
// once and sweeps its band range NBB bands at a time, so the tables are read
// once per workgroup instead of once per band.  Entries of a point are summed
// in lma order, as the reference accumulates them.
// ---------------------------------------------------------------------------
template <class T, int NBB>
__global__ void __launch_bounds__(256)
k_nl_apply(T *__restrict__ out, long ML, int nb, int bands_per_group,
           const int *__restrict__ cov_pts, const int *__restrict__ cov_ptr, int ncov,
           const int *__restrict__ clma, const T *__restrict__ cval,
           const T *__restrict__ coefC, const int *__restrict__ band_active)
{
    extern __shared__ unsigned char smem_raw[];
    const int p0 = blockIdx.x * 256;
    const int npts = min(256, ncov - p0);
    const int e0 = cov_ptr[p0], e1 = cov_ptr[p0 + npts];
    const int ne = e1 - e0;
    int *Llma = reinterpret_cast<int *>(smem_raw);
    T *Lval = reinterpret_cast<T *>(smem_raw + (((size_t)ne * sizeof(int) + 15) & ~(size_t)15));
    for (int e = threadIdx.x; e < ne; e += 256) {
        Llma[e] = clma[e0 + e];
        Lval[e] = cval[e0 + e];
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= npts) return;
    const long pt = cov_pts[p0 + t];
    const int le0 = cov_ptr[p0 + t] - e0, le1 = cov_ptr[p0 + t + 1] - e0;
    const int bg0 = blockIdx.y * bands_per_group;
    const int bg1 = min(nb, bg0 + bands_per_group);
    for (int b0 = bg0; b0 < bg1; b0 += NBB) {
        T acc[NBB];
#pragma unroll
        for (int k = 0; k < NBB; ++k) acc[k] = T(0);
        for (int e = le0; e < le1; ++e) {
            const T v = Lval[e];
            const T *cr = coefC + (long)Llma[e] * nb + b0;
#pragma unroll
            for (int k = 0; k < NBB; ++k)
                if (b0 + k < bg1) acc[k] = acc[k] + v * cr[k];
        }
#pragma unroll
        for (int k = 0; k < NBB; ++k) {
            const int b = b0 + k;
            if (b < bg1 && (!band_active || band_active[b])) {
                T *o = out + (long)b * ML + pt;
                *o = *o + acc[k];
            }
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

template <class T, int MD>
static int launch_hpsi_md(const T *in, T *out, int nb, int flags, const int *band_active,
                          const T *vloc)
{
    Context &c = ctx();
    const GridDesc &g = c.g;
    const StencilGeom &G = sizeof(T) == 8 ? c.geom_h : c.geom_h32;
    constexpr int PX = sizeof(T) == 8 ? 2 : 4;
    StencilCoef<T, MD> cf;
    cf.c0 = (T)g.coef_lap0;
    for (int m = 0; m < MD; ++m) {
        cf.cx[m] = (T)g.coef_lap[0][m];
        cf.cy[m] = (T)g.coef_lap[1][m];
        cf.cz[m] = (T)g.coef_lap[2][m];
    }
    const int tpr = g.N1 / PX;
    const size_t lds = (size_t)(G.ty + 2 * MD) * (g.N1 + 2 * MD) * sizeof(T);
    const long nblk = (long)G.tiles * nb;
    hipLaunchKernelGGL((k_hpsi<T, MD, PX, 2>), dim3((nblk + 7) / 8 * 8), dim3(G.threads), lds,
                       c.stream, in, out, g.ML, g.N1, g.N2, g.N3, G.ty, G.tz, tpr, G.tiles, nb, cf,
                       vloc, flags, band_active);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
static int launch_stencil(const T *in, T *out, int nb, int flags, const int *band_active)
{
    Context &c = ctx();
    const T *vloc = sizeof(T) == 8 ? (const T *)c.vloc64.p : (const T *)c.vloc32.p;
    if ((flags & kLoc) && !vloc) return set_error(kErrState, "local potential not set");
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

template <class T>
int nl_project(const T *in, int nb, T *coef, const int *band_active)
{
    Context &c = ctx();
    if (c.nl.nz <= 0) return kOK;
    constexpr int NBB = 8;
    const long nblk = (long)c.nl.nz * ((nb + NBB - 1) / NBB);
    const T *val = sizeof(T) == 8 ? (const T *)c.nl.row_val64.p : (const T *)c.nl.row_val32.p;
    hipLaunchKernelGGL((k_project<T, NBB>), dim3(nblk), dim3(256), 0, c.stream, in, c.g.ML, nb,
                       c.nl.nz, c.nl.row_ptr.as<int>(), c.nl.row_idx.as<int>(), val,
                       c.nl.scale64.as<double>(), coef, band_active);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int nl_apply(T *out, int nb, const T *coef, const int *band_active)
{
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    if (nl.nz <= 0 || nl.ncov <= 0) return kOK;
    constexpr int NBB = sizeof(T) == 8 ? 8 : 16;
    const int nblk = (nl.ncov + 255) / 256;
    // enough workgroups for 256 CUs: split the bands into groups
    int ngroups = std::max(1, std::min((nb + NBB - 1) / NBB, (2048 + nblk - 1) / nblk));
    int bpg = (nb + ngroups - 1) / ngroups;
    bpg = (bpg + NBB - 1) / NBB * NBB;
    ngroups = (nb + bpg - 1) / bpg;
    const size_t lds = (((size_t)nl.max_blk_entries * sizeof(int) + 15) & ~(size_t)15) +
                       (size_t)nl.max_blk_entries * sizeof(T);
    const T *cval = sizeof(T) == 8 ? (const T *)nl.col_val64.p : (const T *)nl.col_val32.p;
    hipLaunchKernelGGL((k_nl_apply<T, NBB>), dim3(nblk, ngroups), dim3(256), lds, c.stream, out,
                       c.g.ML, nb, bpg, nl.cov_pts.as<int>(), nl.cov_ptr.as<int>(), nl.ncov,
                       nl.col_lma.as<int>(), cval, coef, band_active);
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int hpsi_apply(const T *in, T *out, int nb, int flags, const int *band_active)
{
    Context &c = ctx();
    if (nb <= 0) return kOK;
    if (c.nl.nz <= 0) flags &= ~kNl;
    const int sflags = flags & (kKin | kLoc | kAcc);
    if (flags & kNl) {
        RS_TRY(c.coefC.ensure((size_t)nb * c.nl.nz * sizeof(T)));
        RS_TRY(nl_project<T>(in, nb, c.coefC.as<T>(), band_active));
    }
    if (sflags & (kKin | kLoc)) {
        RS_TRY(launch_stencil<T>(in, out, nb, sflags, band_active));
    } else if (!(flags & kAcc)) {
        RS_HIP(hipMemsetAsync(out, 0, sizeof(T) * (size_t)c.g.ML * nb, c.stream));
    }
    if (flags & kNl) RS_TRY(nl_apply<T>(out, nb, c.coefC.as<T>(), band_active));
    c.hpsi_bands += nb;
    return kOK;
}

template <class T>
int hpsi_stencil_only(const T *in, T *out, int nb)
{
    return launch_stencil<T>(in, out, nb, kKin | kLoc, nullptr);
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

template int hpsi_apply<double>(const double *, double *, int, int, const int *);
template int hpsi_apply<float>(const float *, float *, int, int, const int *);
template int hpsi_stencil_only<double>(const double *, double *, int);
template int hpsi_stencil_only<float>(const float *, float *, int);
template int nl_project<double>(const double *, int, double *, const int *);
template int nl_apply<double>(double *, int, const double *, const int *);
template int nl_apply<float>(float *, int, const float *, const int *);
template int nl_project<float>(const float *, int, float *, const int *);
template int pc_stencil<double>(const double *, double *, double *, double *, double *, int,
                                int, const int *, double *);
template int pc_stencil<float>(const float *, float *, float *, float *, float *, int, int,
                               const int *, double *);

}  // namespace rsdft
