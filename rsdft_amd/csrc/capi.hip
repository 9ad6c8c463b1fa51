// C-ABI entry points (include/rsdft_gpu.h): context, tables, residency and
// the drivers that replace hamiltonian / op_* / conjugate_gradient /
// gram_schmidt / subspace_diag / calc_sweep / calc_density /
// calc_force_ps_nloc2 of the reference.
#include "common.h"
#include "rsdft_gpu.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace rsdft {

static char g_err[1024] = "";

int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *last_error() { return g_err; }

int DevMem::ensure(size_t nbytes)
{
    if (nbytes <= bytes && p) return kOK;
    release();
    if (nbytes == 0) return kOK;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        return set_error(kErrHip, "hipMalloc(%zu) -> %s", nbytes, hipGetErrorString(e));
    }
    bytes = nbytes;
    return kOK;
}

// The context lives on the heap and is never destroyed: a file-scope object
// would free device memory from its destructor during static destruction,
// after the HIP runtime (or a profiler attached to it) has shut down.
// rsdft_gpu_finalize releases everything while HIP is still up.
static Context *const g_ctx = new Context();
Context &ctx() { return *g_ctx; }

// ---- timers ---------------------------------------------------------------
int PhaseEvents::begin(int s, int *handle)
{
    hipEvent_t e0, e1;
    RS_HIP(hipEventCreate(&e0));
    RS_HIP(hipEventCreate(&e1));
    RS_HIP(hipEventRecord(e0, ctx().stream));
    a.push_back(e0);
    b.push_back(e1);
    slot.push_back(s);
    *handle = (int)a.size() - 1;
    return kOK;
}

int PhaseEvents::end(int handle)
{
    RS_HIP(hipEventRecord(b[handle], ctx().stream));
    return kOK;
}

int PhaseEvents::flush()
{
    for (size_t k = 0; k < a.size(); ++k) {
        float ms = 0.0f;
        RS_HIP(hipEventElapsedTime(&ms, a[k], b[k]));
        ctx().timers.ms[slot[k]] += ms;
        (void)hipEventDestroy(a[k]);
        (void)hipEventDestroy(b[k]);
    }
    a.clear();
    b.clear();
    slot.clear();
    return kOK;
}

int phase_begin(int s, int *handle)
{
    *handle = -1;
    return ctx().pe ? ctx().pe->begin(s, handle) : kOK;
}

int phase_end(int handle)
{
    return (ctx().pe && handle >= 0) ? ctx().pe->end(handle) : kOK;
}

// Makes `pe` the timing target of the drivers for the lifetime of the scope.
struct TimedScope {
    PhaseEvents pe;
    TimedScope() { ctx().pe = &pe; }
    ~TimedScope() {
        ctx().pe = nullptr;
        for (size_t k = 0; k < pe.a.size(); ++k) {   // left by an error return
            (void)hipEventDestroy(pe.a[k]);
            (void)hipEventDestroy(pe.b[k]);
        }
    }
};

__global__ void k_count_active(const int *__restrict__ act, int nb, unsigned long long *count)
{
    __shared__ int tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    int n = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) n += act[b] != 0;
    atomicAdd(&tot, n);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(count, (unsigned long long)tot);
}

int count_active_bands(const int *band_active, int nb)
{
    Context &c = ctx();
    if (!c.hpsi_count.p) {
        RS_TRY(c.hpsi_count.ensure(sizeof(unsigned long long)));
        RS_HIP(hipMemsetAsync(c.hpsi_count.p, 0, sizeof(unsigned long long), c.stream));
    }
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, c.stream, band_active, nb,
                       c.hpsi_count.as<unsigned long long>());
    RS_HIP(hipGetLastError());
    return kOK;
}

// ---- scratch of the allocation-free entry point -----------------------------
// Device memory of rsdft_gpu_hamiltonian_noalloc: the projection coefficients
// [lma][band] in FP64.  Every GPU holds its own copy at its own address until
// the process ends; host code reaches it only through the address looked up
// per device at rsdft_gpu_init (the array's name is a host-side stand-in).
constexpr size_t kScratchElems = size_t(2) << 20;   // 16 MiB of FP64
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchElems];
static double *g_scratch_ptr[kScratchMaxDevices] = {};

static int bind_scratch(int device)
{
    if (device < 0 || device >= kScratchMaxDevices)
        return set_error(kErrState, "device %d beyond the %d devices the scratch table holds", device,
                         kScratchMaxDevices);
    if (g_scratch_ptr[device]) return kOK;
    void *p = nullptr;
    RS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double *>(p);
    return kOK;
}

// ---- helpers --------------------------------------------------------------
static int need_init()
{
    if (!ctx().init) return set_error(kErrNotInit, "rsdft_gpu_init has not been called");
    return kOK;
}

static size_t elem_bytes(int prec) { return prec == 0 ? 8 : 4; }

static int ensure_wf_buffers(int MB)
{
    Context &c = ctx();
    const size_t n = (size_t)c.g.ML * MB * elem_bytes(c.prec);
    DevMem *bufs[] = {&c.psi, &c.hx, &c.hp, &c.gk, &c.pgk, &c.pk, &c.rk, &c.ppc, &c.qpc, &c.tmp};
    if (c.wf_prec != c.prec) {
        for (DevMem *b : bufs) b->release();
    }
    for (DevMem *b : bufs) RS_TRY(b->ensure(n));
    RS_TRY(c.dots.ensure(sizeof(double) * 4 * (size_t)MB));
    RS_TRY(c.cg.ensure(MB));
    c.wf_prec = c.prec;
    c.MB = MB;
    return kOK;
}

// host FP64 -> device mirror c.psi (working precision)
static int upload_psi(const double *unk, int MB)
{
    Context &c = ctx();
    const size_t n = (size_t)c.g.ML * MB;
    if (c.prec == 0) {
        RS_HIP(hipMemcpyAsync(c.psi.p, unk, n * 8, hipMemcpyHostToDevice, c.stream));
    } else {
        RS_TRY(c.stage.ensure(n * 8));
        RS_HIP(hipMemcpyAsync(c.stage.p, unk, n * 8, hipMemcpyHostToDevice, c.stream));
        RS_TRY(convert_from_f64<float>(c.stage.as<double>(), c.psi.as<float>(), (long)n));
    }
    return kOK;
}

static int download_psi(double *unk, int MB)
{
    Context &c = ctx();
    const size_t n = (size_t)c.g.ML * MB;
    if (c.prec == 0) {
        RS_HIP(hipMemcpyAsync(unk, c.psi.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    } else {
        RS_TRY(c.stage.ensure(n * 8));
        RS_TRY(convert_to_f64<float>(c.psi.as<float>(), c.stage.as<double>(), (long)n));
        RS_HIP(hipMemcpyAsync(unk, c.stage.p, n * 8, hipMemcpyDeviceToHost, c.stream));
    }
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

// Make c.psi hold unk(ML0, MB): either it is the attached array (nothing to
// do) or it is copied in.  *staged tells the caller to copy back.
static int bind_wf(const double *unk, int ML0, int MB, bool *staged)
{
    Context &c = ctx();
    if (ML0 != c.g.ML) return set_error(kErrArg, "ML0=%d but the grid has %ld points", ML0, c.g.ML);
    if (MB <= 0) return set_error(kErrArg, "MB=%d", MB);
    if (unk == c.host_unk && MB == c.MB && c.wf_prec == c.prec) {
        *staged = false;
        return kOK;
    }
    c.host_unk = nullptr;
    RS_TRY(ensure_wf_buffers(MB));
    RS_TRY(upload_psi(unk, MB));
    *staged = true;
    return kOK;
}

static int check_solver_flags()
{
    Context &c = ctx();
    int info[4] = {0, 0, 0, 0};
    RS_HIP(hipMemcpy(info, c.einfo.p, sizeof(info), hipMemcpyDeviceToHost));
    if (info[2] != 0) return set_error(kErrSolver, "Cholesky of the overlap failed (info=%d)", info[2]);
    if (info[3] != 0) return set_error(kErrSolver, "subspace eigensolve failed (info=%d)", info[3]);
    return kOK;
}

static int clear_solver_flags()
{
    Context &c = ctx();
    RS_TRY(c.einfo.ensure(sizeof(int) * 4));
    RS_HIP(hipMemsetAsync(c.einfo.p, 0, sizeof(int) * 4, c.stream));
    return kOK;
}

template <class T> static T *psiT() { return ctx().psi.as<T>(); }

template <class T>
static int run_sweeps(int MB, int nsweep, double *esp_dev, double *res_dev)
{
    for (int it = 0; it < nsweep; ++it) {
        int hs, hp;
        RS_TRY(phase_begin(kTSweep, &hs));
        RS_TRY(phase_begin(kTCg, &hp));
        RS_TRY(cg_batched<T>(psiT<T>(), MB, esp_dev, res_dev));
        RS_TRY(phase_end(hp));
        if (ctx().sweep_mode == 1) {
            // CG leaves H x in c.hx except for bands with a rejected step
            // (there H x moved and x did not): recompute those bands only
            T *HX = ctx().hx.as<T>();
            RS_TRY(phase_begin(kTGs, &hp));
            RS_TRY(hpsi_apply<T>(psiT<T>(), HX, MB, kKin | kLoc | kNl, ctx().cg.rejected));
            RS_TRY(cholesky_inverse_dev<T>(psiT<T>(), MB));
            RS_TRY(phase_end(hp));
            RS_TRY(phase_begin(kTSd, &hp));
            RS_TRY(rayleigh_ritz_factored_dev<T>(psiT<T>(), HX, MB, esp_dev));
            RS_TRY(phase_end(hp));
        } else {
            RS_TRY(phase_begin(kTGs, &hp));
            RS_TRY(gram_schmidt_dev<T>(psiT<T>(), MB));
            RS_TRY(phase_end(hp));
            RS_TRY(phase_begin(kTSd, &hp));
            RS_TRY(subspace_diag_dev<T>(psiT<T>(), MB, esp_dev));
            RS_TRY(phase_end(hp));
        }
        RS_TRY(phase_end(hs));
    }
    return kOK;
}

}  // namespace rsdft

using namespace rsdft;

extern "C" {

const char *rsdft_gpu_last_error(void) { return last_error(); }

int rsdft_gpu_init(int N1, int N2, int N3, const double *Hgrid, double dV, int Md,
                   const double *coef_lap, double coef_lap0)
{
    Context &c = ctx();
    if (N1 <= 0 || N2 <= 0 || N3 <= 0 || !Hgrid || !coef_lap)
        return set_error(kErrArg, "bad grid arguments");
    if (Md < 1 || Md > kMaxMd) return set_error(kErrArg, "Md=%d out of range", Md);
    if (N1 % 4) return set_error(kErrArg, "N1=%d must be a multiple of 4", N1);
    if (N1 < 2 * Md + 1 || N2 < 2 * Md + 1 || N3 < 2 * Md + 1)
        return set_error(kErrArg, "grid smaller than the stencil");
    if (!c.init) {
        RS_HIP(hipGetDevice(&c.device));
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c.device) ==
                hipSuccess && lds > 0)
            c.max_lds = (size_t)lds;
        RS_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        RS_BLAS(rocblas_create_handle(&c.blas));
        RS_BLAS(rocblas_set_stream(c.blas, c.stream));
        RS_BLAS(rocblas_set_atomics_mode(c.blas, rocblas_atomics_not_allowed));
    }
    RS_TRY(bind_scratch(c.device));
    c.refine_skip = c.refine_fails = 0;
    c.g = GridDesc();
    c.g.N1 = N1;
    c.g.N2 = N2;
    c.g.N3 = N3;
    c.g.Md = Md;
    c.g.ML = (long)N1 * N2 * N3;
    for (int d = 0; d < 3; ++d) c.g.h[d] = Hgrid[d];
    c.g.dV = dV;
    c.g.coef_lap0 = coef_lap0;
    for (int m = 0; m < Md; ++m)
        for (int d = 0; d < 3; ++d) c.g.coef_lap[d][m] = coef_lap[d + 3 * m];
    c.geom_h = choose_hpsi_geom(N1, N2, N3, Md, kHpsiPx64, Md, 8, kHpsiThreads64);
    c.geom_h32 = choose_hpsi_geom(N1, N2, N3, Md, kHpsiPx32, (Md + kHpsiPx32 - 1) / kHpsiPx32 * kHpsiPx32,
                                  4, kHpsiThreads32);
    c.geom_pc = choose_geom(N1, N2, N3, 1, 4, 8);
    if (c.geom_h.threads == 0 || c.geom_h32.threads == 0 || c.geom_pc.threads == 0)
        return set_error(kErrArg, "no stencil tiling for grid %dx%dx%d", N1, N2, N3);
    c.host_unk = nullptr;
    c.MB = 0;
    c.init = true;
    return kOK;
}

int rsdft_gpu_set_vloc(const double *vloc, int s)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (s != 1) return set_error(kErrArg, "spin index %d: one spin channel only", s);
    const size_t n = (size_t)c.g.ML;
    RS_TRY(c.vloc64.ensure(n * 8));
    RS_TRY(c.vloc32.ensure(n * 4));
    RS_HIP(hipMemcpyAsync(c.vloc64.p, vloc, n * 8, hipMemcpyHostToDevice, c.stream));
    RS_TRY(convert_from_f64<float>(c.vloc64.as<double>(), c.vloc32.as<float>(), (long)n));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int rsdft_gpu_set_projectors(int nzlma, int maxmjj, const int *MJJ, const int *JJP,
                             const double *uVk, const int *iuV)
{
    RS_TRY(need_init());
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    const long ML = c.g.ML;
    nl.nz = 0;
    nl.ng = 0;
    nl.natom = 0;   // derivative tables belong to the previous projectors
    if (nzlma <= 0) return kOK;
    // per projector: (grid point, uV) sorted by point, repeated points merged
    // (a sphere larger than the cell lists a point once per image)
    using Entry = std::pair<int, double>;
    std::vector<std::vector<Entry>> pts(nzlma);
    for (int l = 0; l < nzlma; ++l) {
        if (MJJ[l] < 0 || MJJ[l] > maxmjj) return set_error(kErrArg, "MJJ(%d)=%d", l + 1, MJJ[l]);
        auto &P = pts[l];
        P.reserve(MJJ[l]);
        for (int j = 0; j < MJJ[l]; ++j) {
            const int gi = JJP[(long)l * maxmjj + j] - 1;
            if (gi < 0 || gi >= ML)
                return set_error(kErrArg, "JJP(%d,%d)=%d out of grid", j + 1, l + 1, gi + 1);
            P.push_back({gi, uVk[(long)l * maxmjj + j]});
        }
        std::stable_sort(P.begin(), P.end(),
                         [](const Entry &a, const Entry &b) { return a.first < b.first; });
        size_t w = 0;
        for (size_t k = 0; k < P.size(); ++k) {
            if (w > 0 && P[w - 1].first == P[k].first) P[w - 1].second += P[k].second;
            else P[w++] = P[k];
        }
        P.resize(w);
    }
    // groups: consecutive projectors whose point sets overlap by >= 3/4
    std::vector<std::vector<int>> gmem, gset;
    for (int l = 0; l < nzlma; ++l) {
        if (!gmem.empty() && (int)gmem.back().size() < kGroupMax) {
            const std::vector<int> &S = gset.back();
            const auto &P = pts[l];
            size_t a = 0, b = 0, inter = 0;
            while (a < S.size() && b < P.size()) {
                if (S[a] < P[b].first) ++a;
                else if (S[a] > P[b].first) ++b;
                else { ++inter; ++a; ++b; }
            }
            const size_t mx = std::max(S.size(), P.size());
            if (mx > 0 && 4 * inter >= 3 * mx) {
                std::vector<int> U;
                U.reserve(S.size() + P.size() - inter);
                a = 0; b = 0;
                while (a < S.size() || b < P.size()) {
                    if (b == P.size() || (a < S.size() && S[a] < P[b].first)) U.push_back(S[a++]);
                    else if (a == S.size() || S[a] > P[b].first) U.push_back(P[b++].first);
                    else { U.push_back(S[a]); ++a; ++b; }
                }
                gset.back().swap(U);
                gmem.back().push_back(l);
                continue;
            }
        }
        gmem.push_back({l});
        std::vector<int> S(pts[l].size());
        for (size_t k = 0; k < S.size(); ++k) S[k] = pts[l][k].first;
        gset.push_back(std::move(S));
    }
    const int ng = (int)gmem.size();
    int gp = 1;
    for (const auto &m : gmem)
        while (gp < (int)m.size()) gp *= 2;
    std::vector<int> gptr(ng + 1, 0), glma((size_t)ng * gp, -1);
    for (int g = 0; g < ng; ++g) {
        gptr[g + 1] = gptr[g] + (int)gset[g].size();
        for (size_t s2 = 0; s2 < gmem[g].size(); ++s2) glma[(size_t)g * gp + s2] = gmem[g][s2];
    }
    const long npts = gptr[ng];
    std::vector<int> gidx(npts);
    std::vector<double> gval((size_t)npts * gp, 0.0);
    for (int g = 0; g < ng; ++g) {
        const auto &S = gset[g];
        std::copy(S.begin(), S.end(), gidx.begin() + gptr[g]);
        for (size_t s2 = 0; s2 < gmem[g].size(); ++s2) {
            const auto &P = pts[gmem[g][s2]];
            size_t a = 0;
            for (const Entry &e : P) {
                while (S[a] < e.first) ++a;
                gval[((size_t)gptr[g] + a) * gp + s2] = e.second;
            }
        }
    }
    // greedy colouring: groups sharing a grid point get different colours
    std::vector<uint32_t> pcol(ML, 0u);
    std::vector<int> gcol(ng);
    int ncol = 0;
    for (int g = 0; g < ng; ++g) {
        uint32_t used = 0;
        for (int k = gptr[g]; k < gptr[g + 1]; ++k) used |= pcol[gidx[k]];
        int col = 0;
        while (col < 32 && (used >> col) & 1u) ++col;
        if (col == 32) return set_error(kErrArg, "projector spheres overlap too densely");
        gcol[g] = col;
        ncol = std::max(ncol, col + 1);
        for (int k = gptr[g]; k < gptr[g + 1]; ++k) pcol[gidx[k]] |= 1u << col;
    }
    std::vector<int> order;
    order.reserve(ng);
    nl.color_ptr.assign(1, 0);
    for (int col = 0; col < ncol; ++col) {
        for (int g = 0; g < ng; ++g)
            if (gcol[g] == col) order.push_back(g);
        nl.color_ptr.push_back((int)order.size());
    }
    // grid-order fill tables: groups touching each range of kFillRange points
    // (in group order) and, per grid point, its memberships as range-local slots
    const int nrange = (int)((ML + kFillRange - 1) / kFillRange);
    std::vector<int> rptr(nrange + 1, 0), rlast(nrange, -1), rgrp;
    std::vector<std::vector<int>> rlist(nrange);
    for (int g = 0; g < ng; ++g)
        for (int k = gptr[g]; k < gptr[g + 1]; ++k) {
            const int r = gidx[k] / kFillRange;
            if (rlast[r] != g) { rlast[r] = g; rlist[r].push_back(g); }
        }
    int gmax = 0;
    for (int r = 0; r < nrange; ++r) {
        rptr[r + 1] = rptr[r] + (int)rlist[r].size();
        gmax = std::max(gmax, (int)rlist[r].size());
        rgrp.insert(rgrp.end(), rlist[r].begin(), rlist[r].end());
    }
    std::vector<uint32_t> meta(ML, 0u);
    int mmax = 0;
    for (int g = 0; g < ng; ++g)
        for (int k = gptr[g]; k < gptr[g + 1]; ++k) {
            const int i = gidx[k];
            const int n = (int)(meta[i] & 7u);
            mmax = std::max(mmax, n + 1);
            if (n >= kFillMaxM) continue;   // layout rejected below
            const auto &L = rlist[i / kFillRange];
            const int slot = (int)(std::find(L.begin(), L.end(), g) - L.begin());
            meta[i] = (meta[i] & ~7u) | (uint32_t)(n + 1) | ((uint32_t)slot << (3 + 6 * n));
        }
    int fill_mm = 0;
    if (mmax <= kFillMaxM && gmax <= kFillMaxG && ML % 4 == 0 && npts > 0) {
        fill_mm = 1;
        while (fill_mm < mmax) fill_mm *= 2;
        if (fill_mm * gp > 16) fill_mm = 0;   // per-thread uV registers
    }
    std::vector<double> fuv64;
    if (fill_mm) {
        fuv64.assign((size_t)fill_mm * ML * gp, 0.0);
        std::vector<uint8_t> seen(ML, 0);
        for (int g = 0; g < ng; ++g)
            for (int k = gptr[g]; k < gptr[g + 1]; ++k) {
                const int i = gidx[k], m = seen[i]++;
                for (int s2 = 0; s2 < gp; ++s2)
                    fuv64[((size_t)m * ML + i) * gp + s2] = gval[(size_t)k * gp + s2];
            }
    }
    std::vector<float> fuv(fuv64.begin(), fuv64.end());
    std::vector<float> gval32(gval.begin(), gval.end());
    std::vector<double> scale(nzlma);
    for (int l = 0; l < nzlma; ++l) scale[l] = (double)iuV[l] * c.g.dV;
    RS_TRY(nl.g_lma.ensure(sizeof(int) * glma.size()));
    RS_TRY(nl.g_ptr.ensure(sizeof(int) * (ng + 1)));
    RS_TRY(nl.g_idx.ensure(sizeof(int) * std::max<long>(npts, 1)));
    RS_TRY(nl.g_val64.ensure(sizeof(double) * std::max<size_t>(gval.size(), 1)));
    RS_TRY(nl.g_val32.ensure(sizeof(float) * std::max<size_t>(gval.size(), 1)));
    RS_TRY(nl.g_order.ensure(sizeof(int) * ng));
    RS_TRY(nl.scale64.ensure(sizeof(double) * nzlma));
    RS_HIP(hipMemcpy(nl.g_lma.p, glma.data(), sizeof(int) * glma.size(), hipMemcpyHostToDevice));
    RS_HIP(hipMemcpy(nl.g_ptr.p, gptr.data(), sizeof(int) * (ng + 1), hipMemcpyHostToDevice));
    if (npts) {
        RS_HIP(hipMemcpy(nl.g_idx.p, gidx.data(), sizeof(int) * npts, hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.g_val64.p, gval.data(), sizeof(double) * gval.size(), hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.g_val32.p, gval32.data(), sizeof(float) * gval.size(), hipMemcpyHostToDevice));
    }
    RS_HIP(hipMemcpy(nl.g_order.p, order.data(), sizeof(int) * ng, hipMemcpyHostToDevice));
    nl.fill_mm = fill_mm;
    nl.fill_gmax = gmax;
    nl.nrange = nrange;
    if (fill_mm) {
        RS_TRY(nl.fill_rptr.ensure(sizeof(int) * (nrange + 1)));
        RS_TRY(nl.fill_rgrp.ensure(sizeof(int) * std::max<size_t>(rgrp.size(), 1)));
        RS_TRY(nl.fill_meta.ensure(sizeof(uint32_t) * ML));
        RS_TRY(nl.fill_uv.ensure(sizeof(float) * fuv.size()));
        RS_TRY(nl.fill_uv64.ensure(sizeof(double) * fuv64.size()));
        RS_HIP(hipMemcpy(nl.fill_rptr.p, rptr.data(), sizeof(int) * (nrange + 1), hipMemcpyHostToDevice));
        if (!rgrp.empty())
            RS_HIP(hipMemcpy(nl.fill_rgrp.p, rgrp.data(), sizeof(int) * rgrp.size(), hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.fill_meta.p, meta.data(), sizeof(uint32_t) * ML, hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.fill_uv.p, fuv.data(), sizeof(float) * fuv.size(), hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.fill_uv64.p, fuv64.data(), sizeof(double) * fuv64.size(),
                         hipMemcpyHostToDevice));
    }
    RS_HIP(hipMemcpy(nl.scale64.p, scale.data(), sizeof(double) * nzlma, hipMemcpyHostToDevice));
    nl.nz = nzlma;
    nl.ng = ng;
    nl.gp = gp;
    nl.npts = npts;
    return kOK;
}

int rsdft_gpu_set_projector_derivatives(int natom, const int *amap, int mmjj, const int *MJJ_MAP,
                                        const int *JJ_MAP, const double *duVdR)
{
    RS_TRY(need_init());
    Context &c = ctx();
    NonlocalTables &nl = c.nl;
    const int nz = nl.nz;
    if (nz <= 0) return set_error(kErrState, "set projectors first");
    if (natom <= 0) return set_error(kErrArg, "natom=%d", natom);
    std::vector<int> dptr(nz + 1, 0);
    for (int l = 0; l < nz; ++l) {
        if (MJJ_MAP[l] < 0 || MJJ_MAP[l] > mmjj) return set_error(kErrArg, "MJJ_MAP(%d)", l + 1);
        dptr[l + 1] = dptr[l] + MJJ_MAP[l];
    }
    const long nnz = dptr[nz];
    std::vector<int> didx(nnz);
    std::vector<double> dval(3 * nnz);
    for (int l = 0; l < nz; ++l)
        for (int j = 0; j < MJJ_MAP[l]; ++j) {
            const long e = dptr[l] + j;
            const int gi = JJ_MAP[(long)l * mmjj + j] - 1;
            if (gi < 0 || gi >= c.g.ML) return set_error(kErrArg, "JJ_MAP out of grid");
            didx[e] = gi;
            for (int d = 0; d < 3; ++d) dval[3 * e + d] = duVdR[3 * ((long)l * mmjj + j) + d];
        }
    // atom -> projector CSR: [aptr(natom+1) | alma(nz)]
    std::vector<int> cnt(natom + 1, 0), tab(natom + 1 + nz);
    for (int l = 0; l < nz; ++l) {
        const int a = amap[l] - 1;
        if (a < 0 || a >= natom) return set_error(kErrArg, "amap(%d)=%d", l + 1, amap[l]);
        cnt[a + 1]++;
    }
    for (int a = 0; a < natom; ++a) cnt[a + 1] += cnt[a];
    for (int a = 0; a <= natom; ++a) tab[a] = cnt[a];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int l = 0; l < nz; ++l) tab[natom + 1 + fill[amap[l] - 1]++] = l;
    nl.natom = natom;
    nl.nnzd = nnz;
    RS_TRY(nl.d_ptr.ensure(sizeof(int) * (nz + 1)));
    RS_TRY(nl.d_idx.ensure(sizeof(int) * std::max<long>(nnz, 1)));
    RS_TRY(nl.d_val.ensure(sizeof(double) * 3 * std::max<long>(nnz, 1)));
    RS_TRY(nl.amap_dev.ensure(sizeof(int) * tab.size()));
    RS_HIP(hipMemcpy(nl.d_ptr.p, dptr.data(), sizeof(int) * (nz + 1), hipMemcpyHostToDevice));
    if (nnz) {
        RS_HIP(hipMemcpy(nl.d_idx.p, didx.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
        RS_HIP(hipMemcpy(nl.d_val.p, dval.data(), sizeof(double) * 3 * nnz, hipMemcpyHostToDevice));
    }
    RS_HIP(hipMemcpy(nl.amap_dev.p, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
    return kOK;
}

int rsdft_gpu_set_precision(int mode)
{
    RS_TRY(need_init());
    if (mode < 0 || mode > 2) return set_error(kErrArg, "precision mode %d", mode);
    Context &c = ctx();
    if (c.host_unk && c.wf_prec >= 0 && elem_bytes(mode) != elem_bytes(c.wf_prec))
        return set_error(kErrState, "detach (after wf_sync) before changing the storage precision");
    c.prec = mode;
    if (c.wf_prec >= 0 && elem_bytes(mode) == elem_bytes(c.wf_prec)) c.wf_prec = mode;
    return kOK;
}

int rsdft_gpu_set_precond_mode(int mode)
{
    if (mode < 0 || mode > 1) return set_error(kErrArg, "preconditioner mode %d", mode);
    ctx().pc_mode = mode;
    return kOK;
}

int rsdft_gpu_set_eigensolver(int mode)
{
    if (mode < 0 || mode > 3) return set_error(kErrArg, "eigensolver mode %d", mode);
    ctx().eig_mode = mode;
    ctx().refine_skip = ctx().refine_fails = 0;
    return kOK;
}

int rsdft_gpu_set_sweep_mode(int mode)
{
    if (mode < 0 || mode > 1) return set_error(kErrArg, "sweep mode %d", mode);
    ctx().sweep_mode = mode;
    return kOK;
}

int rsdft_gpu_set_solver_params(int ncg, int mloop)
{
    if (ncg < 0 || mloop < 0) return set_error(kErrArg, "ncg=%d mloop=%d", ncg, mloop);
    ctx().ncg = ncg;
    ctx().mloop = mloop;
    return kOK;
}

static int op_entry(int k, int s, const double *tpsi, double *htpsi, int n1, int n2, int ib1,
                    int ib2, int flags)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (k != 1) return set_error(kErrArg, "k=%d: Gamma point only", k);
    if (s != 1) return set_error(kErrArg, "s=%d: one spin channel only", s);
    const long ML0 = (long)n2 - n1 + 1;
    const int nb = ib2 - ib1 + 1;
    if (ML0 != c.g.ML) return set_error(kErrArg, "n2-n1+1=%ld but the grid has %ld points", ML0, c.g.ML);
    if (nb <= 0) return kOK;
    const size_t n = (size_t)ML0 * nb;
    RS_TRY(c.stage.ensure(2 * n * 8));
    double *din = c.stage.as<double>(), *dout = din + n;
    RS_HIP(hipMemcpyAsync(din, tpsi, n * 8, hipMemcpyHostToDevice, c.stream));
    if (flags & kAcc) RS_HIP(hipMemcpyAsync(dout, htpsi, n * 8, hipMemcpyHostToDevice, c.stream));
    if (c.prec == 0) {
        RS_TRY(hpsi_apply<double>(din, dout, nb, flags));
    } else {
        RS_TRY(c.opbuf.ensure(2 * n * 4));
        float *fin = c.opbuf.as<float>(), *fout = fin + n;
        RS_TRY(convert_from_f64<float>(din, fin, (long)n));
        if (flags & kAcc) RS_TRY(convert_from_f64<float>(dout, fout, (long)n));
        RS_TRY(hpsi_apply<float>(fin, fout, nb, flags));
        RS_TRY(convert_to_f64<float>(fout, dout, (long)n));
    }
    RS_HIP(hipMemcpyAsync(htpsi, dout, n * 8, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int rsdft_gpu_hamiltonian(int k, int s, const double *tpsi, double *htpsi, int n1, int n2,
                          int ib1, int ib2)
{
    return op_entry(k, s, tpsi, htpsi, n1, n2, ib1, ib2, kKin | kLoc | kNl);
}

int rsdft_gpu_hamiltonian_noalloc(int k, int s, const double *d_tpsi, double *d_htpsi, int n1,
                                  int n2, int ib1, int ib2)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (k != 1) return set_error(kErrArg, "k=%d: Gamma point only", k);
    if (s != 1) return set_error(kErrArg, "s=%d: one spin channel only", s);
    const long ML0 = (long)n2 - n1 + 1;
    const int nb = ib2 - ib1 + 1;
    if (ML0 != c.g.ML) return set_error(kErrArg, "n2-n1+1=%ld but the grid has %ld points", ML0, c.g.ML);
    if (nb <= 0) return kOK;
    if (!d_tpsi || !d_htpsi) return set_error(kErrArg, "null device array");
    const size_t need = (size_t)nb * std::max(c.nl.nz, 0);
    if (need > kScratchElems)
        return set_error(kErrArg,
                         "%d bands x %d projectors = %zu coefficients exceed the %zu of the "
                         "allocation-free scratch", nb, c.nl.nz, need, kScratchElems);
    double *coef = g_scratch_ptr[c.device];
    if (!coef) return set_error(kErrState, "scratch not bound on device %d", c.device);
    RS_TRY(hpsi_apply<double>(d_tpsi, d_htpsi, nb, kKin | kLoc | kNl, nullptr, coef));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int rsdft_gpu_op_kinetic(int k, const double *tpsi, double *htpsi, int n1, int n2, int ib1, int ib2)
{
    return op_entry(k, 1, tpsi, htpsi, n1, n2, ib1, ib2, kKin);
}

int rsdft_gpu_op_localpot(int s, int mm, int nn, const double *f, double *vf)
{
    return op_entry(1, s, f, vf, 1, mm, 1, nn, kLoc | kAcc);
}

int rsdft_gpu_op_nonlocal(int k, int s, const double *tpsi, double *htpsi, int n1, int n2,
                          int ib1, int ib2)
{
    return op_entry(k, s, tpsi, htpsi, n1, n2, ib1, ib2, kNl | kAcc);
}

int rsdft_gpu_wf_attach(double *unk, int ML0, int MB)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!unk) return set_error(kErrArg, "null unk");
    if (ML0 != c.g.ML) return set_error(kErrArg, "ML0=%d but the grid has %ld points", ML0, c.g.ML);
    RS_TRY(ensure_wf_buffers(MB));
    RS_TRY(upload_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    c.host_unk = unk;
    return kOK;
}

int rsdft_gpu_wf_sync(int dir)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!c.host_unk) return set_error(kErrState, "no wavefunction attached");
    if (dir == 0) return download_psi(c.host_unk, c.MB);
    RS_TRY(upload_psi(c.host_unk, c.MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int rsdft_gpu_wf_detach(void)
{
    ctx().host_unk = nullptr;
    return kOK;
}

int rsdft_gpu_conjugate_gradient(int n1, int n2, int MB, int k, int s, double *unk, double *esp,
                                 double *res)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (k != 1 || s != 1) return set_error(kErrArg, "Gamma point, one spin only");
    bool staged = false;
    RS_TRY(bind_wf(unk, n2 - n1 + 1, MB, &staged));
    TimedScope ts;
    double *ed = c.dots.as<double>();
    int h;
    RS_TRY(phase_begin(kTCg, &h));
    if (c.prec == 0) RS_TRY(cg_batched<double>(psiT<double>(), MB, ed, ed + MB));
    else RS_TRY(cg_batched<float>(psiT<float>(), MB, ed, ed + MB));
    RS_TRY(phase_end(h));
    RS_HIP(hipMemcpyAsync(esp, ed, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipMemcpyAsync(res, ed + MB, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    if (staged) RS_TRY(download_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    return ts.pe.flush();
}

int rsdft_gpu_gram_schmidt(double *unk, int ML0, int MB, int n0, int n1)
{
    (void)n0;
    (void)n1;
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    RS_TRY(clear_solver_flags());
    TimedScope ts;
    int h;
    RS_TRY(phase_begin(kTGs, &h));
    if (c.prec == 0) RS_TRY(gram_schmidt_dev<double>(psiT<double>(), MB));
    else RS_TRY(gram_schmidt_dev<float>(psiT<float>(), MB));
    RS_TRY(phase_end(h));
    if (staged) RS_TRY(download_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    RS_TRY(ts.pe.flush());
    return check_solver_flags();
}

int rsdft_gpu_subspace_diag(double *unk, double *esp, int ML0, int MB)
{
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    RS_TRY(clear_solver_flags());
    TimedScope ts;
    double *ed = c.dots.as<double>();
    int h;
    RS_TRY(phase_begin(kTSd, &h));
    if (c.prec == 0) RS_TRY(subspace_diag_dev<double>(psiT<double>(), MB, ed));
    else RS_TRY(subspace_diag_dev<float>(psiT<float>(), MB, ed));
    RS_TRY(phase_end(h));
    RS_HIP(hipMemcpyAsync(esp, ed, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    if (staged) RS_TRY(download_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    RS_TRY(ts.pe.flush());
    return check_solver_flags();
}

int rsdft_gpu_sweep(double *unk, double *esp, double *res, int ML0, int MB, int nsweep)
{
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    RS_TRY(clear_solver_flags());
    TimedScope ts;
    double *ed = c.dots.as<double>();
    if (c.prec == 0) RS_TRY(run_sweeps<double>(MB, nsweep, ed, ed + MB));
    else RS_TRY(run_sweeps<float>(MB, nsweep, ed, ed + MB));
    if (esp) RS_HIP(hipMemcpyAsync(esp, ed, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    if (res) RS_HIP(hipMemcpyAsync(res, ed + MB, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    if (staged) RS_TRY(download_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    RS_TRY(ts.pe.flush());
    return check_solver_flags();
}

int rsdft_gpu_calc_hartree(const double *rho, double *vh, double *E_h)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!rho || !vh) return set_error(kErrArg, "null rho or vh");
    const size_t n = (size_t)c.g.ML;
    RS_TRY(c.stage.ensure(2 * n * sizeof(double)));
    double *rd = c.stage.as<double>(), *vd = rd + n;
    RS_HIP(hipMemcpyAsync(rd, rho, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    double eh = 0.0;
    RS_TRY(hartree_dev(rd, vd, &eh));
    RS_HIP(hipMemcpyAsync(vh, vd, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    if (E_h) *E_h = eh;
    return kOK;
}

int rsdft_gpu_calc_xc(const double *rho, double *vxc, double *E_xc)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!rho || !vxc) return set_error(kErrArg, "null rho or vxc");
    const size_t n = (size_t)c.g.ML;
    RS_TRY(c.stage.ensure(2 * n * sizeof(double)));
    double *rd = c.stage.as<double>(), *vd = rd + n;
    RS_HIP(hipMemcpyAsync(rd, rho, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    double e3[3];
    RS_TRY(xc_dev(rd, vd, e3));
    RS_HIP(hipMemcpyAsync(vxc, vd, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    if (E_xc) for (int k = 0; k < 3; ++k) E_xc[k] = e3[k];
    return kOK;
}

int rsdft_gpu_calc_fermi(const double *esp, int nb, double nelectron, double ekbt, int kinteg,
                         double *occ, double *efermi)
{
    if (!esp || !occ || !efermi) return set_error(kErrArg, "null argument");
    return calc_fermi_host(esp, nb, nelectron, ekbt, kinteg, occ, efermi);
}

int rsdft_gpu_scf_init(const double *vion, const double *rho, double nelectron, int mmix,
                       double beta, double ekbt, int kinteg)
{
    RS_TRY(need_init());
    return scf_init(vion, rho, nelectron, mmix, beta, ekbt, kinteg);
}

int rsdft_gpu_scf_iterate(double *unk, int ML0, int MB, double *esp, double *occ, double *res,
                          double *info)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!scf_ready()) return set_error(kErrState, "rsdft_gpu_scf_init has not been called");
    if (!esp || !occ || !info) return set_error(kErrArg, "null esp, occ or info");
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    RS_TRY(clear_solver_flags());
    TimedScope ts;
    double *ed = c.dots.as<double>();
    if (c.prec == 0) RS_TRY(scf_iterate_dev<double>(MB, ed, ed + MB, esp, occ, info));
    else RS_TRY(scf_iterate_dev<float>(MB, ed, ed + MB, esp, occ, info));
    if (res) RS_HIP(hipMemcpyAsync(res, ed + MB, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
    if (staged) RS_TRY(download_psi(unk, MB));
    RS_HIP(hipStreamSynchronize(c.stream));
    RS_TRY(ts.pe.flush());
    return check_solver_flags();
}

int rsdft_gpu_scf_get_fields(double *rho, double *vloc, double *vh, double *vxc)
{
    RS_TRY(need_init());
    if (!scf_ready()) return set_error(kErrState, "rsdft_gpu_scf_init has not been called");
    return scf_fields(rho, vloc, vh, vxc);
}

int rsdft_gpu_calc_density(const double *unk, int ML0, int MB, const double *occ, double *rho)
{
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    RS_TRY(c.stage.ensure(sizeof(double) * (c.g.ML + MB)));
    double *rd = c.stage.as<double>(), *od = rd + c.g.ML;
    RS_HIP(hipMemcpyAsync(od, occ, sizeof(double) * MB, hipMemcpyHostToDevice, c.stream));
    if (c.prec == 0) RS_TRY(density_dev<double>(psiT<double>(), MB, od, rd));
    else RS_TRY(density_dev<float>(psiT<float>(), MB, od, rd));
    RS_HIP(hipMemcpyAsync(rho, rd, sizeof(double) * c.g.ML, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    if (staged) c.host_unk = nullptr;
    return kOK;
}

int rsdft_gpu_force_ps_nloc2(const double *unk, int ML0, int MB, const double *occ, double *force2)
{
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    const int na = c.nl.natom;
    if (na <= 0) return set_error(kErrState, "projector derivative tables not set");
    RS_TRY(c.dots.ensure(sizeof(double) * ((size_t)4 * MB + 3 * na + MB)));
    double *fd = c.dots.as<double>() + 4 * (size_t)MB;
    double *od = fd + 3 * na;
    RS_HIP(hipMemcpyAsync(od, occ, sizeof(double) * MB, hipMemcpyHostToDevice, c.stream));
    if (c.prec == 0) RS_TRY(force_dev<double>(psiT<double>(), MB, od, fd));
    else RS_TRY(force_dev<float>(psiT<float>(), MB, od, fd));
    RS_HIP(hipMemcpyAsync(force2, fd, sizeof(double) * 3 * na, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int rsdft_gpu_band_energy_grad(const double *unk, int ML0, int MB, const double *occ, double *E,
                               double *dEdV, double *dEdR)
{
    RS_TRY(need_init());
    Context &c = ctx();
    bool staged = false;
    RS_TRY(bind_wf(unk, ML0, MB, &staged));
    const long ML = c.g.ML;
    if (E) {
        std::vector<double> be(MB);
        double *bd = c.dots.as<double>();
        if (c.prec == 0) {
            RS_TRY(hpsi_apply<double>(psiT<double>(), c.hp.as<double>(), MB, kKin | kLoc | kNl));
            const double *a[1] = {psiT<double>()}, *b[1] = {c.hp.as<double>()};
            RS_TRY(band_dots<double>(1, a, b, MB, c.g.dV, bd, nullptr));
        } else {
            RS_TRY(hpsi_apply<float>(psiT<float>(), c.hp.as<float>(), MB, kKin | kLoc | kNl));
            const float *a[1] = {psiT<float>()}, *b[1] = {c.hp.as<float>()};
            RS_TRY(band_dots<float>(1, a, b, MB, c.g.dV, bd, nullptr));
        }
        RS_HIP(hipMemcpyAsync(be.data(), bd, sizeof(double) * MB, hipMemcpyDeviceToHost, c.stream));
        RS_HIP(hipStreamSynchronize(c.stream));
        double e = 0.0;
        for (int n = 0; n < MB; ++n) e += occ[n] * be[n];
        *E = e;
    }
    if (dEdV) {
        RS_TRY(rsdft_gpu_calc_density(c.host_unk ? c.host_unk : unk, ML0, MB, occ, dEdV));
        for (long i = 0; i < ML; ++i) dEdV[i] *= c.g.dV;
    }
    if (dEdR) {
        RS_TRY(rsdft_gpu_force_ps_nloc2(c.host_unk ? c.host_unk : unk, ML0, MB, occ, dEdR));
        for (int k = 0; k < 3 * c.nl.natom; ++k) dEdR[k] = -dEdR[k];
    }
    if (staged) c.host_unk = nullptr;
    return kOK;
}

int rsdft_gpu_test_hpsi(int nrep, double *times_ms)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!c.host_unk) return set_error(kErrState, "attach wavefunctions first");
    // [0] hamiltonian, [1] op_kinetic, [2] op_localpot, [3] op_nonlocal,
    // [4] stencil kernel as in the hamiltonian, [5] projection kernel,
    // [6] nonlocal term stage, [7] plain copy of all bands
    const int flags[8] = {kKin | kLoc | kNl, kKin, kLoc | kAcc, kNl | kAcc, -1, -2, -3, -4};
    hipEvent_t a, b;
    RS_HIP(hipEventCreate(&a));
    RS_HIP(hipEventCreate(&b));
    const long saved = c.hpsi_bands;
    RS_TRY(c.coefC.ensure((size_t)c.MB * std::max(c.nl.nz, 1) * elem_bytes(c.prec)));
    for (int t = 0; t < 8; ++t) {
        for (int rep = -1; rep < nrep; ++rep) {
            if (rep == 0) RS_HIP(hipEventRecord(a, c.stream));
            if (c.prec == 0) {
                double *x = psiT<double>(), *y = c.hp.as<double>(), *cf = c.coefC.as<double>();
                if (flags[t] >= 0) RS_TRY(hpsi_apply<double>(x, y, c.MB, flags[t]));
                else if (flags[t] == -1) RS_TRY(hpsi_stencil_only<double>(x, y, c.MB));
                else if (flags[t] == -2) RS_TRY(nl_project<double>(x, c.MB, cf));
                else if (flags[t] == -3) RS_TRY(nl_stage<double>(y, c.MB, cf));
                else RS_TRY(copy_bands<double>(x, y, c.MB));
            } else {
                float *x = psiT<float>(), *y = c.hp.as<float>(), *cf = c.coefC.as<float>();
                if (flags[t] >= 0) RS_TRY(hpsi_apply<float>(x, y, c.MB, flags[t]));
                else if (flags[t] == -1) RS_TRY(hpsi_stencil_only<float>(x, y, c.MB));
                else if (flags[t] == -2) RS_TRY(nl_project<float>(x, c.MB, cf));
                else if (flags[t] == -3) RS_TRY(nl_stage<float>(y, c.MB, cf));
                else RS_TRY(copy_bands<float>(x, y, c.MB));
            }
        }
        RS_HIP(hipEventRecord(b, c.stream));
        RS_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        RS_HIP(hipEventElapsedTime(&ms, a, b));
        if (times_ms) times_ms[t] = nrep > 0 ? ms / nrep : 0.0;
    }
    c.hpsi_bands = saved;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return kOK;
}

int rsdft_gpu_test_gemm(int nrep, double *times_ms)
{
    RS_TRY(need_init());
    Context &c = ctx();
    if (!c.host_unk) return set_error(kErrState, "attach wavefunctions first");
    if (!times_ms) return set_error(kErrArg, "null times");
    return time_gemms(nrep, times_ms);
}

int rsdft_gpu_test_eigh(const double *H, int n, int mode, double *w, double *V, double *ms)
{
    RS_TRY(need_init());
    if (!H || !w || n < 1 || mode < 0 || mode > 3)
        return set_error(kErrArg, "test_eigh: n=%d mode=%d", n, mode);
    Context &c = ctx();
    const size_t nn = (size_t)n * n;
    RS_TRY(c.mat2.ensure(sizeof(double) * nn));
    RS_TRY(c.mat3.ensure(sizeof(double) * nn));
    RS_TRY(c.eval.ensure(sizeof(double) * 2 * (size_t)n));
    RS_TRY(clear_solver_flags());
    double *A = c.mat2.as<double>(), *Vb = c.mat3.as<double>(), *D = c.eval.as<double>(), *Vec;
    RS_HIP(hipMemcpyAsync(A, H, sizeof(double) * nn, hipMemcpyHostToDevice, c.stream));
    hipEvent_t a, b;
    RS_HIP(hipEventCreate(&a));
    RS_HIP(hipEventCreate(&b));
    const int saved = c.eig_mode;
    c.eig_mode = mode;
    const int saved_skip = c.refine_skip, saved_fails = c.refine_fails;
    c.refine_skip = c.refine_fails = 0;   // the test always attempts what mode names
    RS_HIP(hipEventRecord(a, c.stream));
    const int rc = eigh_sub(A, n, Vb, D, &Vec);
    c.eig_mode = saved;
    c.refine_skip = saved_skip;
    c.refine_fails = saved_fails;
    RS_TRY(rc);
    RS_HIP(hipEventRecord(b, c.stream));
    RS_HIP(hipMemcpyAsync(w, D, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    if (V) RS_HIP(hipMemcpyAsync(V, Vec, sizeof(double) * nn, hipMemcpyDeviceToHost, c.stream));
    RS_HIP(hipEventSynchronize(b));
    RS_HIP(hipStreamSynchronize(c.stream));
    float t = 0.0f;
    RS_HIP(hipEventElapsedTime(&t, a, b));
    if (ms) *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return check_solver_flags();
}

int rsdft_gpu_get_timers(double *t, int n)
{
    Context &c = ctx();
    if (n > 10) n = 10;
    for (int k = 0; k < n && k < 6; ++k) t[k] = c.timers.ms[k];
    if (n > 6) {
        unsigned long long masked = 0;
        if (c.hpsi_count.p) {
            RS_HIP(hipStreamSynchronize(c.stream));
            RS_HIP(hipMemcpy(&masked, c.hpsi_count.p, sizeof(masked), hipMemcpyDeviceToHost));
        }
        t[6] = (double)c.hpsi_bands + (double)masked;
    }
    if (n > 7) t[7] = (double)c.eig_refined;
    if (n > 8) {   // bands whose CG step was rejected in the last CG call
        t[8] = 0.0;
        if (c.cg.nb > 0 && c.cg.flags.p) {
            std::vector<int> rej(c.cg.nb);
            RS_HIP(hipStreamSynchronize(c.stream));
            RS_HIP(hipMemcpy(rej.data(), c.cg.rejected, sizeof(int) * c.cg.nb,
                             hipMemcpyDeviceToHost));
            for (int r : rej) t[8] += r != 0;
        }
    }
    if (n > 9) {   // band preconditionings by the recurrence (s-step replay unresolved)
        unsigned long long fb = 0;
        if (c.pc_fallback.p) {
            RS_HIP(hipStreamSynchronize(c.stream));
            RS_HIP(hipMemcpy(&fb, c.pc_fallback.p, sizeof(fb), hipMemcpyDeviceToHost));
        }
        t[9] = (double)fb;
    }
    return kOK;
}

int rsdft_gpu_reset_timers(void)
{
    Context &c = ctx();
    for (double &x : c.timers.ms) x = 0.0;
    c.hpsi_bands = 0;
    c.eig_refined = 0;
    if (c.hpsi_count.p)
        RS_HIP(hipMemsetAsync(c.hpsi_count.p, 0, sizeof(unsigned long long), c.stream));
    if (c.pc_fallback.p)
        RS_HIP(hipMemsetAsync(c.pc_fallback.p, 0, sizeof(unsigned long long), c.stream));
    return kOK;
}

int rsdft_gpu_finalize(void)
{
    Context &c = ctx();
    if (!c.init) return kOK;
    (void)hipStreamSynchronize(c.stream);
    DevMem *all[] = {&c.vloc64, &c.vloc32, &c.psi, &c.hx, &c.hp, &c.gk, &c.pgk, &c.pk, &c.rk,
                     &c.ppc, &c.qpc, &c.tmp, &c.coefC, &c.part, &c.stage, &c.mat, &c.mat2,
                     &c.mat3, &c.mat4, &c.eval, &c.einfo, &c.evwork, &c.matT, &c.dots, &c.gemm_part, &c.opbuf, &c.pc_coef,
                     &c.cg.buf, &c.cg.flags, &c.nl.g_lma, &c.nl.g_ptr, &c.nl.g_idx, &c.nl.g_val64,
                     &c.nl.g_val32, &c.nl.g_order, &c.nl.scale64, &c.nl.d_ptr, &c.nl.d_idx, &c.nl.d_val,
                     &c.nl.fill_rptr, &c.nl.fill_rgrp, &c.nl.fill_meta, &c.nl.fill_uv, &c.nl.fill_uv64,
                     &c.nl.amap_dev, &c.hpsi_count, &c.pc_fallback};
    for (DevMem *m : all) m->release();
    (void)scf_release();
    kpoint_release();
    if (c.blas) rocblas_destroy_handle(c.blas);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.blas = nullptr;
    c.stream = nullptr;
    c.init = false;
    c.host_unk = nullptr;
    c.MB = 0;
    c.wf_prec = -1;
    c.prec = 0;
    c.nl.nz = 0;
    c.nl.ng = 0;
    c.nl.npts = 0;
    c.nl.fill_mm = 0;
    c.nl.fill_gmax = 0;
    c.nl.nrange = 0;
    c.nl.color_ptr.clear();
    c.nl.natom = 0;
    c.pc_mode = 1;
    c.eig_mode = 1;
    c.sweep_mode = 1;
    c.eig_refined = 0;
    c.refine_skip = c.refine_fails = 0;
    for (double &x : c.timers.ms) x = 0.0;
    c.hpsi_bands = 0;
    return kOK;
}

}  // extern "C"
