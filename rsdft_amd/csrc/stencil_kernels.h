// Kernels of the FD hamiltonian stencil (included by hamiltonian.hip).
#pragma once
#include "common.h"

namespace rsdft {

template <class T, int MD>
struct StencilCoef {
    T c0;
    T cx[MD], cy[MD], cz[MD];
};

__device__ __forceinline__ int wrapi(int i, int n) {
    i %= n;
    return i < 0 ? i + n : i;
}

// XCD-aware work mapping: workgroup L runs on XCD L % 8; give each XCD a
// contiguous range of work items so neighbouring items share its L2.
__device__ __forceinline__ long xcd_work_index(long L, long total) {
    if ((total & 7) != 0) return L;
    const long per = total >> 3;
    return (L & 7) * per + (L >> 3);
}

// Prefetch distance (planes) of the z-march: (PF+1) must divide 2*MD so the
// register rings of length PF+1 and 2*MD+1+PF wrap together.
template <int MD>
constexpr int stencil_pf() { return MD % 3 == 0 ? 2 : ((2 * MD) % 4 == 0 ? 3 : 1); }
// LDS x pad: MD rounded up to whole chunks of PX points
template <int MD, int PX>
constexpr int stencil_pad() { return (MD + PX - 1) / PX * PX; }

constexpr int cgcd(int a, int b) { return b == 0 ? a : cgcd(b, a % b); }
constexpr int clcm(int a, int b) { return a / cgcd(a, b) * b; }

template <class T, int PX>
struct alignas(sizeof(T) * PX) Chunk {
    T v[PX];
};

// LDS row pitch of k_hpsi2 in chunks: the tpr chunks of a row and nc pad
// chunks on each side, rounded up so that the pitch is tpr modulo one LDS pass
__host__ __device__ constexpr int hpsi2_pitch(int tpr, int nc, int chunk_bytes)
{
    return tpr + (2 * nc * chunk_bytes + kLdsPassBytes - 1) / kLdsPassBytes * kLdsPassBytes /
                     chunk_bytes;
}

// ---------------------------------------------------------------------------
// Md-order hamiltonian stencil: out = [kin] -1/2 lap(in) + [loc] V*in (+ out if ACC)
//
// A workgroup owns a (ty rows) x (tz planes) tile of one band and marches it
// in z.  Thread t owns the PX consecutive x points of chunk c = t % tpr in row
// r = t / tpr, moved as one vector load/store.  z neighbours come from a
// register queue of 2*MD+1+PF planes; x/y neighbours from a double-buffered
// LDS copy of the current plane (rows y0-MD .. y0+ty+MD, x wrapped into pads),
// so every step needs one barrier.  The queue, the y-halo rows, Vloc and the
// accumulated output are rings indexed by the unrolled step u: a load issued
// PF planes ahead is first read PF steps later.  HM = halo chunks per thread
// (ceil(2*MD / ty)).
// ---------------------------------------------------------------------------
template <class T, int MD, int PX, int HM, int FL, int NT = 512, int HE = 0>
__global__ void __launch_bounds__(NT)
k_hpsi2(const T *__restrict__ in, T *__restrict__ out, long ML, int N1, int N2, int N3, int ty,
        int tz, int tiles, int nbands, StencilCoef<T, MD> cf, const T *__restrict__ vloc,
        const int *__restrict__ band_active)
{
    using C = Chunk<T, PX>;
    constexpr int PF = stencil_pf<MD>();
    constexpr int P = stencil_pad<MD, PX>();
    constexpr int NC = P / PX;
    constexpr int QD = 2 * MD + 1 + PF;
    constexpr int RG = PF + 1;
    static_assert(QD % RG == 0, "ring lengths must wrap together");
    // y-halo rows of plane p are requested at step p - HD; HE puts that at the
    // step where the neighbouring tiles request the same rows as their own
    // (p - MD - PF), so the second request is an L2 hit
    constexpr int HD = HE ? MD + PF : PF;
    constexpr int HR = HD + 1;
    constexpr int UL = clcm(QD, HR);
    constexpr bool KIN = (FL & kKin) != 0, LOC = (FL & kLoc) != 0, ACC = (FL & kAcc) != 0;
    extern __shared__ unsigned char smem_raw[];
    C *L = reinterpret_cast<C *>(smem_raw);
    const int tpr = N1 / PX;
    const int WC = hpsi2_pitch(tpr, NC, (int)sizeof(C));
    const int LP = (ty + 2 * MD) * WC;
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
    const int c = tid - r * tpr;
    const long planeC = (long)tpr * N2;
    const long rowC = (long)(y0 + r) * tpr + c;
    const C *src = reinterpret_cast<const C *>(in + (long)band * ML);
    C *dst = reinterpret_cast<C *>(out + (long)band * ML);
    const C *vsrc = reinterpret_cast<const C *>(vloc);

    const int nhalo = 2 * MD * tpr;
    int h_l[HM], h_g[HM], h_c[HM];
    bool h_on[HM];
#pragma unroll
    for (int k = 0; k < HM; ++k) {
        const int hr = tid + k * nthr;
        h_on[k] = hr < nhalo;
        const int hh = h_on[k] ? hr / tpr : 0;
        const int hc = hr - hh * tpr;
        const int yy = wrapi(hh < MD ? y0 - MD + hh : y0 + ty + hh - MD, N2);
        h_l[k] = (hh < MD ? hh : ty + hh) * WC + NC + hc;
        h_g[k] = yy * tpr + hc;
        h_c[k] = hc;
    }

    C q[QD], hb[HR][HM], vq[RG], dq[RG];
#pragma unroll
    for (int k = 0; k < QD - 1; ++k) q[k] = src[wrapi(z0 - MD + k, N3) * planeC + rowC];
#pragma unroll
    for (int k = 0; k < HD; ++k) {
        const long zp = wrapi(z0 + k, N3) * planeC;
#pragma unroll
        for (int h = 0; h < HM; ++h)
            if (h_on[h]) hb[k][h] = src[zp + h_g[h]];
    }
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const long zp = wrapi(z0 + k, N3) * planeC;
        if (LOC) vq[k] = vsrc[zp + rowC];
        if (ACC && k < tz) dq[k] = dst[zp + rowC];
    }
    const int lrow = (r + MD) * WC + NC + c;
    for (int zb = z0; zb < z0 + tz; zb += UL) {
#pragma unroll
        for (int u = 0; u < UL; ++u) {
            const int z = zb + u;
            if (z >= z0 + tz) break;
            C *Lb = L + ((z - z0) & 1) * LP;
            {   // prefetch: queue plane z+MD+PF; halo of plane z+HD; V and out of plane z+PF
                // (offsets <= 2*MD+1 <= N3: one conditional subtract wraps them)
                const int zq = z + MD + PF - (z + MD + PF >= N3 ? N3 : 0);
                const int zh = z + HD - (z + HD >= N3 ? N3 : 0);
                const int zv = z + PF - (z + PF >= N3 ? N3 : 0);
                q[(u + QD - 1) % QD] = src[zq * planeC + rowC];
#pragma unroll
                for (int h = 0; h < HM; ++h)
                    if (h_on[h]) hb[(u + HD) % HR][h] = src[zh * planeC + h_g[h]];
                if (LOC) vq[(u + PF) % RG] = vsrc[zv * planeC + rowC];
                if (ACC && z + PF < z0 + tz) dq[(u + PF) % RG] = dst[zv * planeC + rowC];
            }
            const C cen = q[(u + MD) % QD];
            Lb[lrow] = cen;
            if (c < NC) Lb[lrow + tpr] = cen;
            if (c >= tpr - NC) Lb[lrow - tpr] = cen;
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                if (!h_on[h]) continue;
                const C v = hb[u % HR][h];
                Lb[h_l[h]] = v;
                if (h_c[h] < NC) Lb[h_l[h] + tpr] = v;
                if (h_c[h] >= tpr - NC) Lb[h_l[h] - tpr] = v;
            }
            __syncthreads();
            C o;
            if (KIN) {
                C xw[2 * NC], yw[2 * MD];
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    xw[j] = Lb[lrow - NC + j];
                    xw[NC + j] = Lb[lrow + 1 + j];
                }
#pragma unroll
                for (int m = 0; m < MD; ++m) {
                    yw[m] = Lb[lrow - (m + 1) * WC];
                    yw[MD + m] = Lb[lrow + (m + 1) * WC];
                }
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    T acc = cf.c0 * cen.v[px];
#pragma unroll
                    for (int m = 1; m <= MD; ++m) {
                        const int lo = px - m + P, hi = px + m - PX;   // window offsets
                        const T xl = px - m >= 0 ? cen.v[px - m] : xw[lo / PX].v[lo % PX];
                        const T xr = px + m < PX ? cen.v[px + m] : xw[NC + hi / PX].v[hi % PX];
                        acc = acc + cf.cx[m - 1] * (xl + xr)
                                  + cf.cy[m - 1] * (yw[m - 1].v[px] + yw[MD + m - 1].v[px])
                                  + cf.cz[m - 1] * (q[(u + MD - m) % QD].v[px] +
                                                    q[(u + MD + m) % QD].v[px]);
                    }
                    o.v[px] = acc;
                }
            } else {
#pragma unroll
                for (int px = 0; px < PX; ++px) o.v[px] = T(0);
            }
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                if (LOC) o.v[px] = o.v[px] + vq[u % RG].v[px] * cen.v[px];
                if (ACC) o.v[px] = dq[u % RG].v[px] + o.v[px];
            }
            dst[(long)z * planeC + rowC] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// Strided variant: thread owns the PX points x = c + k*tpr (k < PX) of row r,
// so lanes touch consecutive words in every scalar LDS and HBM access.  Same
// z-march, rings and single barrier as the chunked kernel above.
// ---------------------------------------------------------------------------
template <class T, int MD, int PX, int HM, int FL, int LAZY = 0>
__global__ void __launch_bounds__(512)
k_hpsi_s(const T *__restrict__ in, T *__restrict__ out, long ML, int N1, int N2, int N3, int ty,
         int tz, int tiles, int nbands, StencilCoef<T, MD> cf, const T *__restrict__ vloc,
         const int *__restrict__ band_active)
{
    constexpr int PF = stencil_pf<MD>();
    constexpr int QD = 2 * MD + 1 + PF;
    constexpr int RG = PF + 1;
    static_assert(QD % RG == 0, "ring lengths must wrap together");
    constexpr bool KIN = (FL & kKin) != 0, LOC = (FL & kLoc) != 0, ACC = (FL & kAcc) != 0;
    extern __shared__ unsigned char smem_raw[];
    T *L = reinterpret_cast<T *>(smem_raw);
    const int tpr = N1 / PX;
    const int W = N1 + 2 * MD;
    const int LP = (ty + 2 * MD) * W;
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
    const int c = tid - r * tpr;
    const long plane = (long)N1 * N2;
    const long rowoff = (long)(y0 + r) * N1 + c;
    const T *src = in + (long)band * ML;
    T *dst = out + (long)band * ML;

    const int nhalo = 2 * MD * tpr;
    int h_l[HM], h_g[HM], h_c[HM];
    bool h_on[HM];
#pragma unroll
    for (int k = 0; k < HM; ++k) {
        const int hr = tid + k * nthr;
        h_on[k] = hr < nhalo;
        const int hh = h_on[k] ? hr / tpr : 0;
        const int hc = hr - hh * tpr;
        const int yy = wrapi(hh < MD ? y0 - MD + hh : y0 + ty + hh - MD, N2);
        h_l[k] = (hh < MD ? hh : ty + hh) * W + MD + hc;
        h_g[k] = yy * N1 + hc;
        h_c[k] = hc;
    }

    T q[QD][PX], hb[RG][HM][PX], vq[RG][PX], dq[RG][PX];
#pragma unroll
    for (int k = 0; k < QD - 1; ++k) {
        const T *p = src + wrapi(z0 - MD + k, N3) * plane + rowoff;
#pragma unroll
        for (int px = 0; px < PX; ++px) q[k][px] = p[px * tpr];
    }
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const long zp = wrapi(z0 + k, N3) * plane;
#pragma unroll
        for (int px = 0; px < PX; ++px) {
#pragma unroll
            for (int h = 0; h < HM; ++h)
                if (h_on[h]) hb[k][h][px] = src[zp + h_g[h] + px * tpr];
            if (LOC) vq[k][px] = vloc[zp + rowoff + px * tpr];
            if (ACC && k < tz) dq[k][px] = dst[zp + rowoff + px * tpr];
        }
    }
    const int lrow = (r + MD) * W + MD + c;
    for (int zb = z0; zb < z0 + tz; zb += QD) {
#pragma unroll
        for (int u = 0; u < QD; ++u) {
            const int z = zb + u;
            if (z >= z0 + tz) break;
            T *Lb = L + ((z - z0) & 1) * LP;
            {
                const int zq = z + MD + PF - (z + MD + PF >= N3 ? N3 : 0);
                const int zv = z + PF - (z + PF >= N3 ? N3 : 0);
                const T *p = src + zq * plane + rowoff;
                const long zp = zv * plane;
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    q[(u + QD - 1) % QD][px] = p[px * tpr];
#pragma unroll
                    for (int h = 0; h < HM; ++h)
                        if (h_on[h]) hb[(u + PF) % RG][h][px] = src[zp + h_g[h] + px * tpr];
                    if (LOC) vq[(u + PF) % RG][px] = vloc[zp + rowoff + px * tpr];
                    if (ACC && z + PF < z0 + tz) dq[(u + PF) % RG][px] = dst[zp + rowoff + px * tpr];
                }
            }
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                const T v = q[(u + MD) % QD][px];
                const int x = c + px * tpr;
                Lb[lrow + px * tpr] = v;
                if (x < MD) Lb[lrow + px * tpr + N1] = v;
                if (x >= N1 - MD) Lb[lrow + px * tpr - N1] = v;
#pragma unroll
                for (int h = 0; h < HM; ++h) {
                    if (!h_on[h]) continue;
                    const T w = hb[u % RG][h][px];
                    const int xh = h_c[h] + px * tpr;
                    Lb[h_l[h] + px * tpr] = w;
                    if (xh < MD) Lb[h_l[h] + px * tpr + N1] = w;
                    if (xh >= N1 - MD) Lb[h_l[h] + px * tpr - N1] = w;
                }
            }
            __syncthreads();
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                const T cen = q[(u + MD) % QD][px];
                const T *lr = Lb + lrow + px * tpr;
                T acc = T(0);
                if (KIN && LAZY) {
                    // neighbours read as they are summed: fewer live registers
                    acc = cf.c0 * cen;
#pragma unroll
                    for (int m = 1; m <= MD; ++m)
                        acc = acc + cf.cx[m - 1] * (lr[-m] + lr[m])
                                  + cf.cy[m - 1] * (lr[-m * W] + lr[m * W])
                                  + cf.cz[m - 1] * (q[(u + MD - m) % QD][px] + q[(u + MD + m) % QD][px]);
                } else if (KIN) {
                    T xw[2 * MD], yw[2 * MD];
#pragma unroll
                    for (int m = 0; m < MD; ++m) {
                        xw[m] = lr[-(m + 1)];
                        xw[MD + m] = lr[m + 1];
                        yw[m] = lr[-(m + 1) * W];
                        yw[MD + m] = lr[(m + 1) * W];
                    }
                    acc = cf.c0 * cen;
#pragma unroll
                    for (int m = 0; m < MD; ++m)
                        acc = acc + cf.cx[m] * (xw[m] + xw[MD + m])
                                  + cf.cy[m] * (yw[m] + yw[MD + m])
                                  + cf.cz[m] * (q[(u + MD - 1 - m) % QD][px] +
                                                q[(u + MD + 1 + m) % QD][px]);
                }
                if (LOC) acc = acc + vq[u % RG][px] * cen;
                if (ACC) acc = dq[u % RG][px] + acc;
                dst[(long)z * plane + rowoff + px * tpr] = acc;
            }
        }
    }
}

}  // namespace rsdft
