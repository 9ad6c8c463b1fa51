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

// ---------------------------------------------------------------------------
// Md-order hamiltonian stencil with fused epilogue
//
// Thread t of a workgroup owns row r = t / tpr of the tile and the PX points
// x = c + tpr*px (c = t % tpr), so consecutive lanes touch consecutive words
// in LDS (no bank conflicts) and in HBM (coalesced).  The z queue holds
// 2*MD+1+PF planes: plane z+MD+PF is requested PF steps before it is needed,
// and the y-halo rows of plane z+1 are requested one step ahead into
// registers, so no load latency sits between two barriers.
// ---------------------------------------------------------------------------
template <class T, int MD, int PX, int PF>
__global__ void __launch_bounds__(512, 4)
k_hpsi(const T *__restrict__ in, T *__restrict__ out, long ML, int N1, int N2, int N3,
       int ty, int tz, int tpr, int tiles, int nbands, StencilCoef<T, MD> cf,
       const T *__restrict__ vloc, int flags, const int *__restrict__ band_active)
{
    constexpr int QD = 2 * MD + 1 + PF;
    constexpr int HMAX = 2;   // halo items per thread (ty >= MD guaranteed by the host)
    extern __shared__ unsigned char smem_raw[];
    T *L = reinterpret_cast<T *>(smem_raw);
    const int W = N1 + 2 * MD;            // LDS row: [MD wrapped | N1 | MD wrapped]
    // tiles of one band go to one XCD back to back, so their shared y/z halo
    // planes are served by that XCD's L2
    const long total = (long)tiles * nbands;
    const long wk = xcd_work_index(blockIdx.x, total);
    if (wk >= total) return;
    const int band = (int)(wk / tiles);
    const int tile = (int)(wk - (long)band * tiles);
    if (band_active && !band_active[band]) return;
    const int nty = N2 / ty;
    const int y0 = (tile % nty) * ty;
    const int z0 = (tile / nty) * tz;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int r = tid / tpr;
    const int cx0 = tid - r * tpr;
    const int yrow = y0 + r;
    const long plane = (long)N1 * N2;
    const T *src = in + (long)band * ML;
    T *dst = out + (long)band * ML;
    const bool kin = flags & kKin, loc = flags & kLoc, acc_out = flags & kAcc;
    const long rowoff = (long)yrow * N1 + cx0;

    // halo items of this thread: (LDS row base, global row offset)
    const int nhalo = 2 * MD * tpr;
    int h_lrow[HMAX], h_goff[HMAX], h_x[HMAX];
    bool h_on[HMAX];
#pragma unroll
    for (int k = 0; k < HMAX; ++k) {
        const int hr = tid + k * nthr;
        h_on[k] = hr < nhalo;
        const int hh = h_on[k] ? hr / tpr : 0;
        const int c = hr - hh * tpr;
        int yy = hh < MD ? y0 - MD + hh : y0 + ty + hh - MD;
        yy = wrapi(yy, N2);
        h_lrow[k] = (hh < MD ? hh : ty + hh) * W + MD + c;
        h_goff[k] = yy * N1 + c;
        h_x[k] = c;
    }

    T q[QD][PX];
#pragma unroll
    for (int k = 0; k < QD - 1; ++k) {
        const int zz = wrapi(z0 - MD + k, N3);
        const T *p = src + zz * plane + rowoff;
#pragma unroll
        for (int px = 0; px < PX; ++px) q[k][px] = p[px * tpr];
    }
    T hb[HMAX][PX];
#pragma unroll
    for (int k = 0; k < HMAX; ++k) {
        if (!h_on[k]) continue;
        const T *p = src + (long)z0 * plane + h_goff[k];
#pragma unroll
        for (int px = 0; px < PX; ++px) hb[k][px] = p[px * tpr];
    }
    T *Lrow = L + (r + MD) * W + MD + cx0;
    for (int z = z0; z < z0 + tz; ++z) {
        {   // request plane z+MD+PF
            int zn = z + MD + PF;
            if (zn >= N3) zn -= N3;
            const T *p = src + zn * plane + rowoff;
#pragma unroll
            for (int px = 0; px < PX; ++px) q[QD - 1][px] = p[px * tpr];
        }
        // stage plane z: own row, the y halo rows held in registers, x pads
#pragma unroll
        for (int px = 0; px < PX; ++px) {
            Lrow[px * tpr] = q[MD][px];
            const int x = cx0 + px * tpr;
            if (x < MD || x >= N1 - MD) Lrow[px * tpr + (x < MD ? N1 : -N1)] = q[MD][px];
        }
#pragma unroll
        for (int k = 0; k < HMAX; ++k) {
            if (!h_on[k]) continue;
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                L[h_lrow[k] + px * tpr] = hb[k][px];
                const int x = h_x[k] + px * tpr;
                if (x < MD || x >= N1 - MD)
                    L[h_lrow[k] + px * tpr + (x < MD ? N1 : -N1)] = hb[k][px];
            }
        }
        __syncthreads();
        // request the halo rows of plane z+1
        {
            int zn = z + 1;
            if (zn >= N3) zn -= N3;
#pragma unroll
            for (int k = 0; k < HMAX; ++k) {
                if (!h_on[k]) continue;
                const T *p = src + zn * plane + h_goff[k];
#pragma unroll
                for (int px = 0; px < PX; ++px) hb[k][px] = p[px * tpr];
            }
        }
#pragma unroll
        for (int px = 0; px < PX; ++px) {
            const T *lr = Lrow + px * tpr;
            T acc = T(0);
            if (kin) {
                acc = cf.c0 * q[MD][px];
#pragma unroll
                for (int m = 1; m <= MD; ++m)
                    acc = acc + cf.cx[m - 1] * (lr[-m] + lr[m])
                              + cf.cy[m - 1] * (lr[-m * W] + lr[m * W])
                              + cf.cz[m - 1] * (q[MD - m][px] + q[MD + m][px]);
            }
            const long gi = (long)z * plane + rowoff + px * tpr;
            if (loc) acc = acc + vloc[gi] * q[MD][px];
            if (acc_out) acc = dst[gi] + acc;
            dst[gi] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < QD - 1; ++k) {
#pragma unroll
            for (int px = 0; px < PX; ++px) q[k][px] = q[k + 1][px];
        }
    }
}

}  // namespace rsdft

namespace rsdft {

// ---------------------------------------------------------------------------
// Software-pipelined z-march.  The z queue, the y-halo rows and Vloc are
// rings indexed by the unrolled step u (compile-time slots, no register
// moves), so a load issued PF planes ahead is first read PF steps later and
// its latency overlaps PF steps of work.  QD = 2*MD+1+PF must be a multiple
// of PF+1 so every ring returns to slot 0 after one unrolled block.
// HM = y-halo rows per thread (ceil(2*MD/ty)).
// ---------------------------------------------------------------------------
template <class T, int MD, int PX, int PF, int HM>
__global__ void __launch_bounds__(512)
k_hpsi_pipe(const T *__restrict__ in, T *__restrict__ out, long ML, int N1, int N2, int N3,
            int ty, int tz, int tpr, int tiles, int nbands, StencilCoef<T, MD> cf,
            const T *__restrict__ vloc, int flags, const int *__restrict__ band_active)
{
    constexpr int QD = 2 * MD + 1 + PF;
    constexpr int RG = PF + 1;
    static_assert(QD % RG == 0, "queue length must be a multiple of the prefetch ring");
    extern __shared__ unsigned char smem_raw[];
    T *L = reinterpret_cast<T *>(smem_raw);
    const int W = N1 + 2 * MD;
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
    const T *src = in + (long)band * ML;
    T *dst = out + (long)band * ML;
    const bool kin = flags & kKin, loc = flags & kLoc, acc_out = flags & kAcc;
    const long rowoff = (long)yrow * N1 + cx0;

    const int nhalo = 2 * MD * tpr;
    int h_lrow[HM], h_goff[HM], h_x[HM];
    bool h_on[HM];
#pragma unroll
    for (int k = 0; k < HM; ++k) {
        const int hr = tid + k * nthr;
        h_on[k] = hr < nhalo;
        const int hh = h_on[k] ? hr / tpr : 0;
        const int c = hr - hh * tpr;
        int yy = hh < MD ? y0 - MD + hh : y0 + ty + hh - MD;
        yy = wrapi(yy, N2);
        h_lrow[k] = (hh < MD ? hh : ty + hh) * W + MD + c;
        h_goff[k] = yy * N1 + c;
        h_x[k] = c;
    }

    T q[QD][PX], hb[RG][HM][PX], vq[RG][PX];
    // prologue: planes z0-MD .. z0+MD+PF-1 into queue slots 0..QD-2; halo/V planes z0..z0+PF-1
#pragma unroll
    for (int k = 0; k < QD - 1; ++k) {
        const T *p = src + (long)wrapi(z0 - MD + k, N3) * plane + rowoff;
#pragma unroll
        for (int px = 0; px < PX; ++px) q[k][px] = p[px * tpr];
    }
#pragma unroll
    for (int k = 0; k < PF; ++k) {
        const long zp = (long)wrapi(z0 + k, N3) * plane;
#pragma unroll
        for (int h = 0; h < HM; ++h) {
            if (!h_on[h]) continue;
#pragma unroll
            for (int px = 0; px < PX; ++px) hb[k][h][px] = src[zp + h_goff[h] + px * tpr];
        }
        if (loc) {
#pragma unroll
            for (int px = 0; px < PX; ++px) vq[k][px] = vloc[zp + rowoff + px * tpr];
        }
    }
    T *Lrow = L + (r + MD) * W + MD + cx0;
    for (int zb = z0; zb < z0 + tz; zb += QD) {
#pragma unroll
        for (int u = 0; u < QD; ++u) {
            const int z = zb + u;
            if (z >= z0 + tz) break;
            // prefetch plane z+MD+PF of the queue, halo rows and V of plane z+PF
            {
                const T *p = src + (long)wrapi(z + MD + PF, N3) * plane + rowoff;
#pragma unroll
                for (int px = 0; px < PX; ++px) q[(u + QD - 1) % QD][px] = p[px * tpr];
                const long zp = (long)wrapi(z + PF, N3) * plane;
#pragma unroll
                for (int h = 0; h < HM; ++h) {
                    if (!h_on[h]) continue;
#pragma unroll
                    for (int px = 0; px < PX; ++px)
                        hb[(u + PF) % RG][h][px] = src[zp + h_goff[h] + px * tpr];
                }
                if (loc) {
#pragma unroll
                    for (int px = 0; px < PX; ++px) vq[(u + PF) % RG][px] = vloc[zp + rowoff + px * tpr];
                }
            }
            // stage plane z
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                const T v = q[(u + MD) % QD][px];
                Lrow[px * tpr] = v;
                const int x = cx0 + px * tpr;
                if (x < MD || x >= N1 - MD) Lrow[px * tpr + (x < MD ? N1 : -N1)] = v;
            }
#pragma unroll
            for (int h = 0; h < HM; ++h) {
                if (!h_on[h]) continue;
#pragma unroll
                for (int px = 0; px < PX; ++px) {
                    const T v = hb[u % RG][h][px];
                    L[h_lrow[h] + px * tpr] = v;
                    const int x = h_x[h] + px * tpr;
                    if (x < MD || x >= N1 - MD) L[h_lrow[h] + px * tpr + (x < MD ? N1 : -N1)] = v;
                }
            }
            __syncthreads();
#pragma unroll
            for (int px = 0; px < PX; ++px) {
                const T *lr = Lrow + px * tpr;
                const T cen = q[(u + MD) % QD][px];
                T acc = T(0);
                if (kin) {
                    acc = cf.c0 * cen;
#pragma unroll
                    for (int m = 1; m <= MD; ++m)
                        acc = acc + cf.cx[m - 1] * (lr[-m] + lr[m])
                                  + cf.cy[m - 1] * (lr[-m * W] + lr[m * W])
                                  + cf.cz[m - 1] * (q[(u + MD - m) % QD][px] + q[(u + MD + m) % QD][px]);
                }
                const long gi = (long)z * plane + rowoff + px * tpr;
                if (loc) acc = acc + vq[u % RG][px] * cen;
                if (acc_out) acc = dst[gi] + acc;
                dst[gi] = acc;
            }
            __syncthreads();
        }
    }
}

}  // namespace rsdft
