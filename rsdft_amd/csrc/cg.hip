// Batched conjugate-gradient eigensolver and its preconditioner.
//
// conjugate_gradient_1 (cg_module.f90:97-430) treats one band at a time and
// the bands never interact (SWGS = 0 makes cggs a no-op,
// cggs_module.f90:23-25), so every band runs the same fixed sequence here in
// lock-step: one launch per step over all bands, per-band FP64 scalars on the
// device, a per-band `active` flag for the early exit of cg_module.f90:218.
// Reductions are two-stage (per-chunk partials, then a fixed-order sum per
// band) so results are deterministic.
#include "common.h"

#include <algorithm>

namespace rsdft {

template <class T>
int pc_stencil(const T *in, T *Q, T *R, T *PP, T *PG, int nb, int mode, const int *active,
               double *part);

constexpr int kChunkThreads = 256;
// elements of one band per chunk: 32 KB of each array (large enough that the
// number of workgroups does not limit the streaming rate)
template <class T> constexpr int kChunk = 32768 / (int)sizeof(T);

template <class T>
static inline int nchunk_of(long ML) { return (int)((ML + kChunk<T> - 1) / kChunk<T>); }

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <int ND>
__device__ __forceinline__ void block_sum_write(double (&s)[ND], double *part, long stride) {
    __shared__ double red[ND][kChunkThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
        const double v = wave_sum(s[k]);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < ND) {
        double t = 0.0;
        for (int w = 0; w < kChunkThreads / 64; ++w) t += red[threadIdx.x][w];
        part[threadIdx.x * stride] = t;
    }
}

template <class T, int ND>
struct PtrSet {
    const T *a[ND];
    const T *b[ND];
};

// part[(k*nb + b)*nchunk + chunk] = sum over the chunk of a_k*b_k (band b)
// 16-byte vectors of T; ML is a multiple of 4 (N1 is), so every band and
// chunk starts 16-byte aligned and holds whole vectors
template <class T> struct Vec16 { typedef T type __attribute__((ext_vector_type(16 / sizeof(T)))); };
template <class T> constexpr int kVW = 16 / (int)sizeof(T);

template <class T, int ND>
__global__ void __launch_bounds__(kChunkThreads)
k_dots(PtrSet<T, ND> p, long ML, int nb, int nchunk, double *part, const int *active)
{
    typedef typename Vec16<T>::type V;
    constexpr int VW = kVW<T>;
    const int b = blockIdx.y, chunk = blockIdx.x;
    if (active && !active[b]) return;
    double s[ND];
#pragma unroll
    for (int k = 0; k < ND; ++k) s[k] = 0.0;
    const long off = (long)b * ML;
    const long i0 = (long)chunk * kChunk<T>;
    const long i1 = min(ML, i0 + kChunk<T>);
    for (long i = i0 + (long)threadIdx.x * VW; i < i1; i += (long)kChunkThreads * VW) {
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const V x = *reinterpret_cast<const V *>(p.a[k] + off + i);
            const V y = *reinterpret_cast<const V *>(p.b[k] + off + i);
#pragma unroll
            for (int e = 0; e < VW; ++e) s[k] += (double)x[e] * (double)y[e];
        }
    }
    block_sum_write<ND>(s, part + ((long)b) * nchunk + chunk, (long)nb * nchunk);
}

// The six inner products of the CG step (cg_module.f90:283-296) from four
// arrays, each read once: (x,x), (p,x), (p,p), (x,Hx), (p,Hx), (p,Hp)
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_dots_cg(const T *__restrict__ X, const T *__restrict__ P, const T *__restrict__ HX,
          const T *__restrict__ HP, long ML, int nb, int nchunk, double *part, const int *active)
{
    typedef typename Vec16<T>::type V;
    constexpr int VW = kVW<T>;
    const int b = blockIdx.y, chunk = blockIdx.x;
    if (active && !active[b]) return;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const long off = (long)b * ML;
    const long i0 = (long)chunk * kChunk<T>;
    const long i1 = min(ML, i0 + kChunk<T>);
    for (long i = i0 + (long)threadIdx.x * VW; i < i1; i += (long)kChunkThreads * VW) {
        const V x = *reinterpret_cast<const V *>(X + off + i);
        const V pv = *reinterpret_cast<const V *>(P + off + i);
        const V hx = *reinterpret_cast<const V *>(HX + off + i);
        const V hp = *reinterpret_cast<const V *>(HP + off + i);
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            const double xe = x[e], pe = pv[e], hxe = hx[e], hpe = hp[e];
            s[0] += xe * xe;
            s[1] += pe * xe;
            s[2] += pe * pe;
            s[3] += xe * hxe;
            s[4] += pe * hxe;
            s[5] += pe * hpe;
        }
    }
    block_sum_write<6>(s, part + ((long)b) * nchunk + chunk, (long)nb * nchunk);
}

// sum of partials in fixed order
__device__ __forceinline__ double sum_parts(const double *part, int nchunk) {
    double t = 0.0;
    for (int c = 0; c < nchunk; ++c) t += part[c];
    return t;
}

// Collapse the per-chunk partial sums of `rows` rows (row r at part[r*nchunk])
// into part[r*nchunk]: one wave per row, lanes take strided chunks, then a
// fixed shuffle tree (deterministic).  The per-band scalar kernels that
// follow read one value per row (nchunk = 1).
__global__ void __launch_bounds__(256) k_collapse_parts(double *part, long rows, int nchunk)
{
    const long r = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    double *p = part + r * nchunk;
    double t = 0.0;
    for (int c = lane; c < nchunk; c += 64) t += p[c];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
    if (lane == 0) p[0] = t;
}

static void collapse_parts(double *part, long rows, int nchunk)
{
    if (nchunk <= 1) return;
    hipLaunchKernelGGL(k_collapse_parts, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0,
                       ctx().stream, part, rows, nchunk);
}

__global__ void k_sum_parts(const double *part, int nb, int nchunk, int nd, double scale,
                            double *out, const int *active)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (active && !active[b]) return;
    for (int k = 0; k < nd; ++k)
        out[k * nb + b] = scale * sum_parts(part + ((long)k * nb + b) * nchunk, nchunk);
}

// ---------------------------------------------------------------------------
// 2x2 generalized eigenproblem as dsygv(1,'V','U') solves it
// (cg_module.f90:318): Cholesky B = U^T U, C = U^-T A U^-1, one Jacobi
// rotation, z = U^-1 y, ascending w, z^T B z = 1.  Column-major 2x2.
// ---------------------------------------------------------------------------
__device__ __host__ inline void sygv2(const double *A, const double *B, double *W, double *Z)
{
    const double u11 = sqrt(B[0]);
    const double u12 = B[2] / u11;
    const double u22 = sqrt(B[3] - u12 * u12);
    const double i11 = 1.0 / u11, i12 = -u12 / (u11 * u22), i22 = 1.0 / u22;
    const double a = A[0], b = A[2], c = A[3];
    const double c11 = i11 * a * i11;
    const double c12 = i11 * (a * i12 + b * i22);
    const double c22 = i12 * (a * i12 + b * i22) + i22 * (b * i12 + c * i22);
    double w1, w2, y11, y21, y12, y22;
    if (c12 == 0.0) {
        w1 = c11; w2 = c22; y11 = 1.0; y21 = 0.0; y12 = 0.0; y22 = 1.0;
    } else {
        const double tau = (c22 - c11) / (2.0 * c12);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        w1 = c11 - t * c12; w2 = c22 + t * c12;
        y11 = cs; y21 = -sn; y12 = sn; y22 = cs;
    }
    if (w2 < w1) {
        double tw = w1; w1 = w2; w2 = tw;
        double t1 = y11, t2 = y21; y11 = y12; y21 = y22; y12 = t1; y22 = t2;
    }
    W[0] = w1; W[1] = w2;
    Z[0] = i11 * y11 + i12 * y21; Z[1] = i22 * y21;
    Z[2] = i11 * y12 + i12 * y22; Z[3] = i22 * y22;
}

// ---------------------------------------------------------------------------
// per-band scalar steps (one thread per band)
// ---------------------------------------------------------------------------
struct CGView {
    double *E, *E1, *rb, *res, *gkgk, *bk, *u1, *u2, *W, *rr0, *rr1, *pa, *pb;
    int *active, *accept, *pc_active, *rejected, *pc_ill, *pc_ok;
};

static CGView view(CGScalars &s) {
    return CGView{s.E, s.E1, s.rb, s.res, s.gkgk, s.bk, s.u1, s.u2, s.W, s.rr0, s.rr1,
                  s.pa, s.pb, s.active, s.accept, s.pc_active, s.rejected, s.pc_ill, s.pc_ok};
}

// E = dV <x|Hx>, E1 = 1e10 (cg_module.f90:162,176-182)
__global__ void k_cg_init(CGView v, const double *part, int nb, int nchunk, double dV)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    v.E[b] = dV * part[(long)b * nchunk];   // collapsed partials
    v.E1[b] = 1.0e10;
    v.active[b] = 1;
    v.rejected[b] = 0;
}

// rb = dV <g|g>
__global__ void k_cg_rb(CGView v, const double *part, int nb, int nchunk, double dV)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.active[b]) return;
    v.rb[b] = dV * part[(long)b * nchunk];
}

// top of the icg loop (cg_module.f90:213-219): res = rb/c1^2, early exits
__global__ void k_cg_top(CGView v, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.active[b]) return;
    v.res[b] = v.rb[b] / 4.0;
    if (v.rb[b] < 0.0) v.active[b] = 0;
    else if (fabs(v.E[b] - v.E1[b]) < 1.0e-15) v.active[b] = 0;
}

// rb = dV <Pg|g>; beta (cg_module.f90:237-261), for the bands mask selects
__global__ void k_cg_pgg(CGView v, const double *part, int nb, int nchunk, double dV, int first,
                         const int *mask)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !mask[b]) return;
    const double rb = dV * part[(long)b * nchunk];
    v.rb[b] = rb;
    v.bk[b] = first ? 0.0 : rb / v.gkgk[b];
    v.gkgk[b] = rb;
}

// the same from (Pg, g) given by the s-step preconditioner's Gram matrix,
// for the bands it resolved
__global__ void k_cg_pgg_gram(CGView v, const double *pgg, int nb, double dV, int first)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.pc_ok[b]) return;
    const double rb = dV * pgg[b];
    v.rb[b] = rb;
    v.bk[b] = first ? 0.0 : rb / v.gkgk[b];
    v.gkgk[b] = rb;
}

// 2x2 Rayleigh-Ritz with the swap rule and the sign fix (cg_module.f90:286-333)
__global__ void k_cg_2x2(CGView v, const double *part, int nb, int nchunk, double dV)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.active[b]) return;
    double w[6];
    for (int k = 0; k < 6; ++k) w[k] = dV * part[((long)k * nb + b) * nchunk];
    const double Bm[4] = {w[0], w[1], w[1], w[2]};
    const double Am[4] = {w[3], w[4], w[4], w[5]};
    double W[2], Z[4];
    sygv2(Am, Bm, W, Z);
    const double E = v.E[b];
    if (fabs(W[0] - E) > 1.0e-1 && fabs(W[1] - E) <= 1.0e-1) {
        Z[0] = Z[2]; Z[1] = Z[3]; W[0] = W[1];
    }
    if (Z[0] < 0.0) { Z[0] = -Z[0]; Z[1] = -Z[1]; }
    v.u1[b] = Z[0];
    v.u2[b] = Z[1];
    v.W[b] = W[0];
    v.E1[b] = E;
    v.E[b] = W[0];
}

// split the active bands by whether the s-step replay resolved them; the
// unresolved ones are counted: count[0] cumulative, count[1] this call
__global__ void k_cg_split_ill(CGView v, int nb, unsigned long long *count)
{
    __shared__ int tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    int n = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        const int a = v.active[b];
        const int ill = a && v.pc_ill[b];
        v.pc_ill[b] = ill;
        v.pc_ok[b] = a && !ill;
        n += ill;
    }
    atomicAdd(&tot, n);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (tot) atomicAdd(count, (unsigned long long)tot);
        count[1] = (unsigned long long)tot;
    }
}

// new residual norm and the reject rule (cg_module.f90:349-364)
__global__ void k_cg_reject(CGView v, const double *part, int nb, int nchunk, double dV)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.active[b]) return;
    const double rb = dV * part[(long)b * nchunk];
    v.rb[b] = rb;
    if (rb / v.res[b] > 1.0e8) {
        v.E[b] = v.E1[b];
        v.accept[b] = 0;
        v.rejected[b] = 1;   // Hx moved, x did not
    } else {
        v.accept[b] = 1;
    }
}

__global__ void k_cg_final(CGView v, int nb, double *esp, double *res)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    esp[b] = v.E[b];
    res[b] = v.res[b];
}

// preconditioner scalars (cgpc_hprsdft_module.f90:165-249), no dV
__global__ void k_pc_init(CGView v, const double *part, int nb, int nchunk, const int *mask)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (!mask[b]) { v.pc_active[b] = 0; return; }
    const double rr1 = sum_parts(part + (long)b * nchunk, nchunk);
    v.rr1[b] = rr1;
    v.pc_active[b] = rr1 < 1.0e-24 ? 0 : 1;
}

__global__ void k_pc_alpha(CGView v, const double *part, int nb, int nchunk)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.pc_active[b]) return;
    const double pAp = sum_parts(part + (long)b * nchunk, nchunk);
    v.rr0[b] = v.rr1[b];
    v.pa[b] = v.rr0[b] / pAp;
}

__global__ void k_pc_beta(CGView v, const double *part, int nb, int nchunk)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb || !v.pc_active[b]) return;
    const double rr1 = sum_parts(part + (long)b * nchunk, nchunk);
    v.rr1[b] = rr1;
    v.pb[b] = rr1 / v.rr0[b];
}

// ---------------------------------------------------------------------------
// element-wise band kernels (grid: chunks x bands)
// ---------------------------------------------------------------------------
#define BAND_LOOP_BEGIN                                                       \
    const int b = blockIdx.y, chunk = blockIdx.x;                             \
    const long off = (long)b * ML;                                            \
    const long i0 = (long)chunk * kChunk<T>;                                \
    const long i1 = min(ML, i0 + kChunk<T>);

// g = -2 (Hx - E x); part = sum g^2
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_cg_grad(const T *HX, const T *X, T *G, CGView v, long ML, int nchunk, double *part)
{
    BAND_LOOP_BEGIN
    if (!v.active[b]) return;
    typedef typename Vec16<T>::type V;
    constexpr int VW = kVW<T>;
    const T E = (T)v.E[b];
    double s[1] = {0.0};
    for (long i = i0 + (long)threadIdx.x * VW; i < i1; i += (long)kChunkThreads * VW) {
        const V hx = *reinterpret_cast<const V *>(HX + off + i);
        const V x = *reinterpret_cast<const V *>(X + off + i);
        V g;
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            g[e] = T(-2) * (hx[e] - E * x[e]);
            s[0] += (double)g[e] * (double)g[e];
        }
        *reinterpret_cast<V *>(G + off + i) = g;
    }
    block_sum_write<1>(s, part + (long)b * nchunk + chunk, 0);
}

// p = Pg (first) or p = Pg + beta p
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_cg_newp(const T *PG, T *P, CGView v, long ML, int first, const int *mask)
{
    BAND_LOOP_BEGIN
    if (!mask[b]) return;
    const T bk = (T)v.bk[b];
    if (first) {
        for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads) P[off + i] = PG[off + i];
    } else {
        for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads)
            P[off + i] = PG[off + i] + bk * P[off + i];
    }
}

// Hx = u1 Hx + u2 Hp; g = -2 (Hx - W (u1 x + u2 p)); part = sum g^2
// (cg_module.f90:337-349).  When u1 > 1/20 the new x = u1 x + u2 p is stored
// here already (the reject rule of cg_module.f90:360-368 almost never fires;
// k_cg_accept undoes it then, losing at most a factor 20 in rounding), which
// saves a pass over x and p.
__device__ __forceinline__ bool cg_speculative(double u1) { return u1 > 0.05; }

// G == null: the gradient is only summed (the last CG step, whose gradient
// no later step reads)
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_cg_update(T *HX, const T *HP, T *X, const T *P, T *G, CGView v, long ML, int nchunk,
            double *part)
{
    BAND_LOOP_BEGIN
    if (!v.active[b]) return;
    typedef typename Vec16<T>::type V;
    constexpr int VW = kVW<T>;
    const T u1 = (T)v.u1[b], u2 = (T)v.u2[b], w = (T)v.W[b];
    const bool spec = cg_speculative(v.u1[b]);
    double s[1] = {0.0};
    for (long i = i0 + (long)threadIdx.x * VW; i < i1; i += (long)kChunkThreads * VW) {
        const V hx0 = *reinterpret_cast<const V *>(HX + off + i);
        const V hp = *reinterpret_cast<const V *>(HP + off + i);
        const V x0 = *reinterpret_cast<const V *>(X + off + i);
        const V pv = *reinterpret_cast<const V *>(P + off + i);
        V hx, xn, g;
#pragma unroll
        for (int e = 0; e < VW; ++e) {
            hx[e] = u1 * hx0[e] + u2 * hp[e];
            xn[e] = u1 * x0[e] + u2 * pv[e];
            g[e] = T(-2) * (hx[e] - w * xn[e]);
            s[0] += (double)g[e] * (double)g[e];
        }
        *reinterpret_cast<V *>(HX + off + i) = hx;
        if (G) *reinterpret_cast<V *>(G + off + i) = g;
        if (spec) *reinterpret_cast<V *>(X + off + i) = xn;
    }
    block_sum_write<1>(s, part + (long)b * nchunk + chunk, 0);
}

// x = u1 x + u2 p unless rejected (cg_module.f90:366-368); for the bands
// k_cg_update already moved, a rejection restores x = (x_new - u2 p) / u1.
// Same chunk x band grid as the other element-wise kernels: almost every
// band has nothing to do here and its workgroups return at once.
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_cg_accept(T *X, const T *P, CGView v, long ML)
{
    BAND_LOOP_BEGIN
    if (!v.active[b]) return;
    const bool spec = cg_speculative(v.u1[b]);
    if (spec == (v.accept[b] != 0)) return;
    typedef typename Vec16<T>::type V;
    constexpr int VW = kVW<T>;
    const T u1 = (T)v.u1[b], u2 = (T)v.u2[b];
    for (long i = i0 + (long)threadIdx.x * VW; i < i1; i += (long)kChunkThreads * VW) {
        V x = *reinterpret_cast<const V *>(X + off + i);
        const V pv = *reinterpret_cast<const V *>(P + off + i);
#pragma unroll
        for (int e = 0; e < VW; ++e) x[e] = spec ? (x[e] - u2 * pv[e]) / u1 : u1 * x[e] + u2 * pv[e];
        *reinterpret_cast<V *>(X + off + i) = x;
    }
}

// Pg += a pp; unless last: r -= a q, part = sum r^2
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_pc_update(T *PG, const T *PP, T *R, const T *Q, CGView v, long ML, int nchunk, int last,
            double *part)
{
    BAND_LOOP_BEGIN
    if (!v.pc_active[b]) return;
    const T a = (T)v.pa[b];
    double s[1] = {0.0};
    if (last) {
        for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads)
            PG[off + i] = PG[off + i] + a * PP[off + i];
        return;
    }
    for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads) {
        PG[off + i] = PG[off + i] + a * PP[off + i];
        const T r = R[off + i] - a * Q[off + i];
        R[off + i] = r;
        s[0] += (double)r * (double)r;
    }
    block_sum_write<1>(s, part + (long)b * nchunk + chunk, 0);
}

// pp = r + b pp
template <class T>
__global__ void __launch_bounds__(kChunkThreads)
k_pc_newdir(const T *R, T *PP, CGView v, long ML)
{
    BAND_LOOP_BEGIN
    if (!v.pc_active[b]) return;
    const T bb = (T)v.pb[b];
    for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads)
        PP[off + i] = R[off + i] + bb * PP[off + i];
}

template <class T>
__global__ void k_copy_active(const T *src, T *dst, long ML, const int *active)
{
    BAND_LOOP_BEGIN
    if (active && !active[b]) return;
    for (long i = i0 + threadIdx.x; i < i1; i += kChunkThreads) dst[off + i] = src[off + i];
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
int CGScalars::ensure(int MB)
{
    RS_TRY(buf.ensure(sizeof(double) * 13 * (size_t)MB));
    RS_TRY(flags.ensure(sizeof(int) * 6 * (size_t)MB));
    double *d = buf.as<double>();
    double **slots[13] = {&E, &E1, &rb, &res, &gkgk, &bk, &u1, &u2, &W, &rr0, &rr1, &pa, &pb};
    for (int k = 0; k < 13; ++k) *slots[k] = d + (size_t)k * MB;
    int *f = flags.as<int>();
    active = f;
    accept = f + MB;
    pc_active = f + 2 * (size_t)MB;
    rejected = f + 3 * (size_t)MB;
    pc_ill = f + 4 * (size_t)MB;
    pc_ok = f + 5 * (size_t)MB;
    nb = MB;
    return kOK;
}

template <class T>
int band_dots(int ndot, const T *const *a, const T *const *b, int nb, double scale,
              double *out, const int *active)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    const int nch = nchunk_of<T>(ML);
    RS_TRY(c.part.ensure(sizeof(double) * (size_t)6 * nb * nch));
    double *part = c.part.as<double>();
    dim3 grid(nch, nb);
    switch (ndot) {
#define RS_DOTS_CASE(N)                                                                      \
    case N: {                                                                                \
        PtrSet<T, N> p;                                                                      \
        for (int k = 0; k < N; ++k) { p.a[k] = a[k]; p.b[k] = b[k]; }                        \
        hipLaunchKernelGGL((k_dots<T, N>), grid, dim3(kChunkThreads), 0, c.stream, p, ML, nb, \
                           nch, part, active);                                               \
        break;                                                                               \
    }
        RS_DOTS_CASE(1)
        RS_DOTS_CASE(2)
        RS_DOTS_CASE(3)
        RS_DOTS_CASE(6)
#undef RS_DOTS_CASE
    default: return set_error(kErrArg, "band_dots: ndot=%d", ndot);
    }
    RS_HIP(hipGetLastError());
    if (out) {
        hipLaunchKernelGGL(k_sum_parts, dim3((nb + 127) / 128), dim3(128), 0, c.stream, part, nb,
                           nch, ndot, scale, out, active);
        RS_HIP(hipGetLastError());
    }
    return kOK;
}

template <class T>
int precond_batched(const T *g, T *pg, int nb, const int *active)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    const int nch = nchunk_of<T>(ML);
    CGView v = view(c.cg);
    const dim3 bgrid((nb + 127) / 128), bblk(128);
    const dim3 egrid(nch, nb), eblk(kChunkThreads);
    if (c.mloop <= 0) {
        hipLaunchKernelGGL((k_copy_active<T>), egrid, eblk, 0, c.stream, g, pg, ML, active);
        RS_HIP(hipGetLastError());
        return kOK;
    }
    T *R = c.rk.as<T>(), *PP = c.ppc.as<T>(), *Q = c.qpc.as<T>();
    const int tiles = c.geom_pc.tiles;
    RS_TRY(c.part.ensure(sizeof(double) * (size_t)6 * nb * (nch > tiles ? nch : tiles)));
    double *part = c.part.as<double>();
    RS_TRY(pc_stencil<T>(g, Q, R, PP, pg, nb, 0, active, part));
    hipLaunchKernelGGL(k_pc_init, bgrid, bblk, 0, c.stream, v, part, nb, tiles, active);
    for (int il = 1; il <= c.mloop; ++il) {
        RS_TRY(pc_stencil<T>(PP, Q, R, PP, pg, nb, 1, c.cg.pc_active, part));
        hipLaunchKernelGGL(k_pc_alpha, bgrid, bblk, 0, c.stream, v, part, nb, tiles);
        const int last = il == c.mloop;
        hipLaunchKernelGGL((k_pc_update<T>), egrid, eblk, 0, c.stream, pg, PP, R, Q, v, ML, nch,
                           last, part);
        if (last) break;
        hipLaunchKernelGGL(k_pc_beta, bgrid, bblk, 0, c.stream, v, part, nb, nch);
        hipLaunchKernelGGL((k_pc_newdir<T>), egrid, eblk, 0, c.stream, R, PP, v, ML);
    }
    RS_HIP(hipGetLastError());
    return kOK;
}

template <class T>
int cg_batched(T *X, int nb, double *esp_dev, double *res_dev)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    const double dV = c.g.dV;
    const int nch = nchunk_of<T>(ML);
    RS_TRY(c.cg.ensure(nb));
    // one size for every reduction of the sweep, so no later ensure() moves it
    const int maxparts = std::max(nch, c.geom_pc.tiles);
    RS_TRY(c.part.ensure(std::max(sizeof(double) * (size_t)6 * nb * maxparts,
                                  sizeof(double) * precond_sstep_part_elems(nb))));
    RS_TRY(c.pc_coef.ensure(sizeof(double) * (size_t)nb * (std::max(c.mloop, 1) + 1)));
    if (!c.pc_fallback.p) {
        RS_TRY(c.pc_fallback.ensure(2 * sizeof(unsigned long long)));
        RS_HIP(hipMemsetAsync(c.pc_fallback.p, 0, 2 * sizeof(unsigned long long), c.stream));
    }
    // the s-step preconditioner hands (Pg, g) and the new direction over directly
    const bool sstep = c.pc_mode == 1 && c.mloop >= 1 && c.mloop <= 4;
    double *part = c.part.as<double>();
    CGView v = view(c.cg);
    T *HX = c.hx.as<T>(), *HP = c.hp.as<T>(), *G = c.gk.as<T>(), *PG = c.pgk.as<T>(),
      *P = c.pk.as<T>();
    const int *act = c.cg.active;
    const dim3 bgrid((nb + 127) / 128), bblk(128);
    const dim3 egrid(nch, nb), eblk(kChunkThreads);

    RS_TRY(hpsi_apply<T>(X, HX, nb, kKin | kLoc | kNl));
    {
        const T *a[1] = {X}, *bb[1] = {HX};
        RS_TRY(band_dots<T>(1, a, bb, nb, 1.0, nullptr, nullptr));
        collapse_parts(part, nb, nch);
        hipLaunchKernelGGL(k_cg_init, bgrid, bblk, 0, c.stream, v, part, nb, nch, dV);
    }
    hipLaunchKernelGGL((k_cg_grad<T>), egrid, eblk, 0, c.stream, HX, X, G, v, ML, nch, part);
    collapse_parts(part, nb, nch);
    hipLaunchKernelGGL(k_cg_rb, bgrid, bblk, 0, c.stream, v, part, nb, nch, dV);
    for (int icg = 1; icg <= c.ncg + 1; ++icg) {
        hipLaunchKernelGGL(k_cg_top, bgrid, bblk, 0, c.stream, v, nb);
        if (icg == c.ncg + 1) break;
        int hpc;
        RS_TRY(phase_begin(kTPc, &hpc));
        // Pg by the step-by-step recurrence, (Pg, g), beta and P = Pg + beta P
        // for the bands `mask` selects
        auto stepwise = [&](const int *mask) -> int {
            RS_TRY(precond_batched<T>(G, PG, nb, mask));
            const T *a[1] = {PG}, *bb[1] = {G};
            RS_TRY(band_dots<T>(1, a, bb, nb, 1.0, nullptr, mask));
            collapse_parts(part, nb, nch);
            hipLaunchKernelGGL(k_cg_pgg, bgrid, bblk, 0, c.stream, v, part, nb, nch, dV,
                               icg == 1 ? 1 : 0, mask);
            hipLaunchKernelGGL((k_cg_newp<T>), egrid, eblk, 0, c.stream, PG, P, v, ML,
                               icg == 1 ? 1 : 0, mask);
            return kOK;
        };
        if (sstep) {
            // Gram pass -> d, (Pg,g) -> rb, beta -> P = Pg + beta P in the apply pass
            // (the preconditioner timer covers both passes and so also the
            // direction update fused into the second); bands whose replay the
            // Gram forms cannot resolve take the recurrence
            double *dcoef = c.pc_coef.as<double>(), *pgg = dcoef + (size_t)nb * c.mloop;
            RS_TRY(precond_sstep_coeffs<T>(G, nb, act, part, dcoef, pgg, v.pc_ill));
            hipLaunchKernelGGL(k_cg_split_ill, dim3(1), dim3(256), 0, c.stream, v, nb,
                               c.pc_fallback.as<unsigned long long>());
            hipLaunchKernelGGL(k_cg_pgg_gram, bgrid, bblk, 0, c.stream, v, pgg, nb, dV,
                               icg == 1 ? 1 : 0);
            RS_TRY(precond_sstep_apply<T>(G, P, nb, v.pc_ok, dcoef, 1, icg == 1 ? nullptr : v.bk));
            // the recurrence's launches only when some band needs them
            unsigned long long nill = 0;
            RS_HIP(hipMemcpyAsync(&nill, c.pc_fallback.as<unsigned long long>() + 1, sizeof(nill),
                                  hipMemcpyDeviceToHost, c.stream));
            RS_HIP(hipStreamSynchronize(c.stream));
            if (nill) RS_TRY(stepwise(v.pc_ill));
            RS_TRY(phase_end(hpc));
        } else {
            RS_TRY(stepwise(act));
            RS_TRY(phase_end(hpc));
        }
        RS_TRY(hpsi_apply<T>(P, HP, nb, kKin | kLoc | kNl, act));
        {
            hipLaunchKernelGGL((k_dots_cg<T>), egrid, eblk, 0, c.stream, X, P, HX, HP, ML, nb, nch,
                               part, act);
            collapse_parts(part, 6L * nb, nch);
            hipLaunchKernelGGL(k_cg_2x2, bgrid, bblk, 0, c.stream, v, part, nb, nch, dV);
        }
        hipLaunchKernelGGL((k_cg_update<T>), egrid, eblk, 0, c.stream, HX, HP, X, P,
                           icg == c.ncg ? (T *)nullptr : G, v, ML, nch, part);
        collapse_parts(part, nb, nch);
        hipLaunchKernelGGL(k_cg_reject, bgrid, bblk, 0, c.stream, v, part, nb, nch, dV);
        hipLaunchKernelGGL((k_cg_accept<T>), egrid, eblk, 0, c.stream, X, P, v, ML);
    }
    hipLaunchKernelGGL(k_cg_final, bgrid, bblk, 0, c.stream, v, nb, esp_dev, res_dev);
    RS_HIP(hipGetLastError());
    return kOK;
}

template int band_dots<double>(int, const double *const *, const double *const *, int, double,
                               double *, const int *);
template int band_dots<float>(int, const float *const *, const float *const *, int, double,
                              double *, const int *);
template int precond_batched<double>(const double *, double *, int, const int *);
template int precond_batched<float>(const float *, float *, int, const int *);
template int cg_batched<double>(double *, int, double *, double *);
template int cg_batched<float>(float *, int, double *, double *);

}  // namespace rsdft
