// The reference's on-disk formats: the wavefunction file of
// simple_wf_io_write / simple_wf_io_read (io_write_module.f90:19-264,
// io_read_module.f90:15-521) and the density/potential file of write_data /
// read_data (io_module.f90:68-249, 254-455), for one process and the
// periodic grid (SYStype = 0, IO_ctrl = 0).
//
// Both are Fortran sequential unformatted files: every record is framed by
// its byte length as a 4-byte little-endian integer before and after the
// data (the convention of gfortran and flang).  Grid values are stored in
// the order of the map LL2(3, ML) written into the file, which for one
// process is the natural order i1 fastest with 0-based indices; a reader
// maps each value through the file's LL2, so files written by another
// decomposition are read correctly too.
//
// The wavefunctions of the attached array are written from the device
// mirror one band at a time (FP64 or FP32 storage), and read straight into
// it, without a host copy of all bands.
#include "common.h"
#include "rsdft_gpu.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace rsdft {

namespace {

class RecordWriter {
  public:
    explicit RecordWriter(const char *path) : f_(std::fopen(path, "wb")) {}
    ~RecordWriter() { if (f_) std::fclose(f_); }
    bool ok() const { return f_ != nullptr && good_; }
    void rec(const void *data, size_t n) {
        if (!f_ || n > 0x7fffffffu) { good_ = false; return; }
        const int32_t len = (int32_t)n;
        good_ = good_ && std::fwrite(&len, 4, 1, f_) == 1;
        if (n) good_ = good_ && std::fwrite(data, 1, n, f_) == n;
        good_ = good_ && std::fwrite(&len, 4, 1, f_) == 1;
    }
    template <class T> void rec(const std::vector<T> &v) { rec(v.data(), v.size() * sizeof(T)); }
    void text(const char *s) {   // character(64), blank padded
        char buf[64];
        std::memset(buf, ' ', 64);
        std::memcpy(buf, s, std::min<size_t>(64, std::strlen(s)));
        rec(buf, 64);
    }
    int close() {
        if (!f_) return 1;
        const int r = std::fclose(f_);
        f_ = nullptr;
        return r != 0 || !good_;
    }

  private:
    FILE *f_;
    bool good_ = true;
};

class RecordReader {
  public:
    explicit RecordReader(const char *path) : f_(std::fopen(path, "rb")) {}
    ~RecordReader() { if (f_) std::fclose(f_); }
    bool open() const { return f_ != nullptr; }
    // the next record into buf (resized); false at end of file or a broken frame
    bool next(std::vector<unsigned char> &buf) {
        int32_t a = 0, b = 0;
        if (!f_ || std::fread(&a, 4, 1, f_) != 1 || a < 0) return false;
        buf.resize((size_t)a);
        if (a && std::fread(buf.data(), 1, (size_t)a, f_) != (size_t)a) return false;
        if (std::fread(&b, 4, 1, f_) != 1 || b != a) return false;
        return true;
    }
    void rewind() { if (f_) std::rewind(f_); }

  private:
    FILE *f_;
};

template <class T>
bool take(const std::vector<unsigned char> &r, size_t count, T *out)
{
    if (r.size() < count * sizeof(T)) return false;
    std::memcpy(out, r.data(), count * sizeof(T));
    return true;
}

struct Grid3 {
    int N1, N2, N3;
    long ML() const { return (long)N1 * N2 * N3; }
};

int check_grid(int N1, int N2, int N3, int ML0)
{
    if (N1 <= 0 || N2 <= 0 || N3 <= 0) return set_error(kErrArg, "grid %dx%dx%d", N1, N2, N3);
    if ((long)N1 * N2 * N3 != ML0)
        return set_error(kErrArg, "ML0=%d but the grid %dx%dx%d has %ld points", ML0, N1, N2, N3,
                         (long)N1 * N2 * N3);
    return kOK;
}

// LL2(3, ML) of the periodic grid on one process (construct_gridmap_sub)
std::vector<int32_t> natural_map(const Grid3 &g)
{
    std::vector<int32_t> L(3 * (size_t)g.ML());
    size_t i = 0;
    for (int i3 = 0; i3 < g.N3; ++i3)
        for (int i2 = 0; i2 < g.N2; ++i2)
            for (int i1 = 0; i1 < g.N1; ++i1) {
                L[i++] = i1;
                L[i++] = i2;
                L[i++] = i3;
            }
    return L;
}

// natural index of every file position from the file's LL2; -1 on a bad map
int file_order(const std::vector<int32_t> &LL, const Grid3 &g, std::vector<long> &pos)
{
    const long ML = g.ML();
    pos.assign(ML, -1);
    std::vector<unsigned char> seen(ML, 0);
    for (long i = 0; i < ML; ++i) {
        const int a = LL[3 * i], b = LL[3 * i + 1], c = LL[3 * i + 2];
        if (a < 0 || a >= g.N1 || b < 0 || b >= g.N2 || c < 0 || c >= g.N3)
            return set_error(kErrArg, "grid map entry %ld = (%d,%d,%d) outside the grid", i + 1, a,
                             b, c);
        const long k = a + (long)g.N1 * (b + (long)g.N2 * c);
        if (seen[k]) return set_error(kErrArg, "grid map lists point (%d,%d,%d) twice", a, b, c);
        seen[k] = 1;
        pos[i] = k;
    }
    return kOK;
}

// the attached device mirror, when unk is the array attached to the context
bool device_backed(const double *unk, int MB)
{
    Context &c = ctx();
    return c.init && unk && unk == c.host_unk && MB == c.MB && c.wf_prec >= 0;
}

// band n of the device mirror as FP64 into host (staged through the device
// scratch band c.opbuf)
int download_band(int n, double *host)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    if (c.wf_prec == 0) {
        RS_HIP(hipMemcpyAsync(host, c.psi.as<double>() + (size_t)n * ML, ML * 8,
                              hipMemcpyDeviceToHost, c.stream));
    } else {
        RS_TRY(c.opbuf.ensure(ML * 8));
        RS_TRY(convert_to_f64<float>(c.psi.as<float>() + (size_t)n * ML, c.opbuf.as<double>(), ML));
        RS_HIP(hipMemcpyAsync(host, c.opbuf.p, ML * 8, hipMemcpyDeviceToHost, c.stream));
    }
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

int upload_band(int n, const double *host)
{
    Context &c = ctx();
    const long ML = c.g.ML;
    if (c.wf_prec == 0) {
        RS_HIP(hipMemcpyAsync(c.psi.as<double>() + (size_t)n * ML, host, ML * 8,
                              hipMemcpyHostToDevice, c.stream));
    } else {
        RS_TRY(c.opbuf.ensure(ML * 8));
        RS_HIP(hipMemcpyAsync(c.opbuf.p, host, ML * 8, hipMemcpyHostToDevice, c.stream));
        RS_TRY(convert_from_f64<float>(c.opbuf.as<double>(), c.psi.as<float>() + (size_t)n * ML, ML));
    }
    RS_HIP(hipStreamSynchronize(c.stream));
    return kOK;
}

}  // namespace

}  // namespace rsdft

using namespace rsdft;

extern "C" {

int rsdft_gpu_write_wf(const char *file, int N1, int N2, int N3, const double *unk, int ML0,
                       int MB, int OC, int MBwr1, int MBwr2, const double *occ, const double *aa,
                       const double *bb)
{
    const Grid3 g{N1, N2, N3};
    RS_TRY(check_grid(N1, N2, N3, ML0));
    if (!file || !unk || !occ) return set_error(kErrArg, "write_wf: null argument");
    if (MB <= 0) return set_error(kErrArg, "write_wf: MB=%d", MB);
    const bool single = OC == 4 || OC == 5;
    if (!(OC == 1 || OC == 3 || OC == 4 || OC == 5 || OC == 11 || OC == 13))
        return set_error(kErrArg, "write_wf: OC=%d writes no wavefunctions", OC);
    // MBwr1/MBwr2 outside 1..MB fall back to all bands (io_write_module.f90:80-81)
    if (MBwr1 < 1 || MBwr2 < MBwr1 || MB < MBwr1) MBwr1 = 1;
    if (MBwr2 < 1 || MBwr2 < MBwr1 || MB < MBwr2) MBwr2 = MB;
    const bool dev = device_backed(unk, MB);
    if (dev && ctx().g.ML != ML0) return set_error(kErrState, "write_wf: attached grid differs");
    RecordWriter w(file);
    if (!w.ok()) return set_error(kErrArg, "write_wf: cannot open %s", file);
    const int32_t ML = (int32_t)ML0;
    const int32_t head[4] = {ML, N1, N2, N3}, bands[3] = {MB, MBwr1, MBwr2};
    if (OC < 10) {
        const int32_t minus1 = -1;
        w.rec(&minus1, 4);
        w.rec(head, sizeof(head));
        w.rec(bands, sizeof(bands));
        const int32_t kms[3] = {1, 1, 1};                // MK, MS, MMBZ: Gamma, one spin
        w.rec(kms, sizeof(kms));
        const int32_t ctl[3] = {0, OC, 1};               // IO_ctrl, OC, TYPE_WF (real)
        w.rec(ctl, sizeof(ctl));
        w.rec(natural_map(g));
        w.rec(occ, sizeof(double) * MB);
        std::vector<double> abk(21, 0.0);                // aa, bb, kbb(3, 1) = 0
        for (int k = 0; k < 9; ++k) {
            abk[k] = aa ? aa[k] : 0.0;
            abk[9 + k] = bb ? bb[k] : 0.0;
        }
        w.rec(abk);
        // one process: np(0:7) = 1, each distribution one block from 0
        const std::vector<int32_t> np(8, 1);
        w.rec(np);
        const int32_t dg[2] = {0, ML}, db[2] = {0, MB}, dk[2] = {0, 1}, ds[2] = {0, 1};
        w.rec(dg, 8);
        w.rec(db, 8);
        w.rec(dk, 8);
        w.rec(ds, 8);
    } else {   // old format (io_write_module.f90:146-151)
        w.rec(head, sizeof(head));
        w.rec(bands, sizeof(bands));
        w.rec(natural_map(g));
        w.rec(occ, sizeof(double) * MB);
    }
    std::vector<double> band(dev ? (size_t)ML0 : 0);
    std::vector<float> sp(single ? (size_t)ML0 : 0);
    for (int n = MBwr1 - 1; n < MBwr2; ++n) {
        const double *src = unk + (size_t)n * ML0;
        if (dev) {
            RS_TRY(download_band(n, band.data()));
            src = band.data();
        }
        if (single) {
            for (long i = 0; i < ML0; ++i) sp[i] = (float)src[i];
            w.rec(sp);
        } else {
            w.rec(src, sizeof(double) * ML0);
        }
    }
    if (w.close()) return set_error(kErrArg, "write_wf: write to %s failed", file);
    return kOK;
}

int rsdft_gpu_read_wf(const char *file, int N1, int N2, int N3, double *unk, int ML0, int MB,
                      double *occ, int *mb_file)
{
    const Grid3 g{N1, N2, N3};
    RS_TRY(check_grid(N1, N2, N3, ML0));
    if (!file || !unk) return set_error(kErrArg, "read_wf: null argument");
    RecordReader r(file);
    if (!r.open()) return set_error(kErrArg, "read_wf: cannot open %s", file);
    std::vector<unsigned char> b;
    int32_t first = 0, head[4], bands[3], ctl[3] = {0, 0, 1}, kms[3] = {1, 1, 1};
    if (!r.next(b) || !take(b, 1, &first)) return set_error(kErrArg, "read_wf: %s: empty", file);
    const bool newfmt = first < 0;
    if (!newfmt) r.rewind();
    bool ok = r.next(b) && take(b, 4, head) && r.next(b) && take(b, 3, bands);
    if (ok && newfmt) ok = r.next(b) && take(b, 3, kms) && r.next(b) && take(b, 3, ctl);
    if (!ok) return set_error(kErrArg, "read_wf: %s: bad header", file);
    if (head[0] != ML0 || head[1] != N1 || head[2] != N2 || head[3] != N3)
        return set_error(kErrArg, "read_wf: file grid %d (%dx%dx%d), this grid %d (%dx%dx%d)",
                         head[0], head[1], head[2], head[3], ML0, N1, N2, N3);
    if (kms[0] != 1 || kms[1] != 1)
        return set_error(kErrArg, "read_wf: %d k-points x %d spins: Gamma, one spin only",
                         kms[0], kms[1]);
    if (newfmt && ctl[2] != 1)
        return set_error(kErrArg, "read_wf: complex wavefunctions (TYPE_WF=%d)", ctl[2]);
    std::vector<int32_t> LL(3 * (size_t)ML0);
    if (!r.next(b) || !take(b, LL.size(), LL.data()))
        return set_error(kErrArg, "read_wf: %s: bad grid map", file);
    std::vector<long> pos;
    RS_TRY(file_order(LL, g, pos));
    const int MBf = bands[0], MB1 = bands[1], MB2 = bands[2];
    std::vector<double> occf((size_t)std::max(MBf, 1));
    if (!r.next(b) || !take(b, (size_t)MBf, occf.data()))
        return set_error(kErrArg, "read_wf: %s: bad occupations", file);
    if (occ)
        for (int n = 0; n < std::min(MB, MBf); ++n) occ[n] = occf[n];
    if (newfmt)
        for (int k = 0; k < 6; ++k)   // aa bb kbb, np, 4 distributions
            if (!r.next(b)) return set_error(kErrArg, "read_wf: %s: bad header", file);
    const bool single = ctl[1] == 4 || ctl[1] == 5;
    const bool dev = device_backed(unk, MB);
    std::vector<double> band((size_t)ML0);
    for (int n = MB1; n <= MB2; ++n) {
        if (!r.next(b)) return set_error(kErrArg, "read_wf: %s: band %d missing", file, n);
        if (n < 1 || n > MB) continue;   // not held here (io_read_module.f90:286-289)
        double *dst = dev ? band.data() : unk + (size_t)(n - 1) * ML0;
        if (single) {
            if (b.size() != 4 * (size_t)ML0) return set_error(kErrArg, "read_wf: band %d size", n);
            const float *v = reinterpret_cast<const float *>(b.data());
            for (long i = 0; i < ML0; ++i) dst[pos[i]] = (double)v[i];
        } else {
            if (b.size() != 8 * (size_t)ML0) return set_error(kErrArg, "read_wf: band %d size", n);
            const double *v = reinterpret_cast<const double *>(b.data());
            for (long i = 0; i < ML0; ++i) dst[pos[i]] = v[i];
        }
        if (dev) {
            std::memcpy(unk + (size_t)(n - 1) * ML0, band.data(), 8 * (size_t)ML0);
            RS_TRY(upload_band(n - 1, band.data()));
        }
    }
    if (mb_file) *mb_file = MBf;
    return kOK;
}

int rsdft_gpu_write_vrho(const char *file, int N1, int N2, int N3, int ML0, int MSP,
                         const double *rho, const double *vloc, const double *vh,
                         const double *vxc, const double *aa, int nelement, int natom,
                         const double *aa_atom, const int *ki_atom, const int *zn_atom,
                         double efermi)
{
    const Grid3 g{N1, N2, N3};
    RS_TRY(check_grid(N1, N2, N3, ML0));
    if (!file || !rho || !vloc || !vh || !vxc || MSP < 1 || MSP > 2)
        return set_error(kErrArg, "write_vrho: bad argument");
    if (natom < 0 || nelement < 0 || (natom && (!aa_atom || !ki_atom)) || (nelement && !zn_atom))
        return set_error(kErrArg, "write_vrho: bad atom arguments");
    RecordWriter w(file);
    if (!w.ok()) return set_error(kErrArg, "write_vrho: cannot open %s", file);
    w.text("version3.0, comment_length=64");
    const int32_t head[4] = {ML0, N1, N2, N3};
    w.rec(head, sizeof(head));
    w.rec(natural_map(g));
    w.text("lattice aa(3,3)");
    std::vector<double> a9(9, 0.0);
    if (aa) std::memcpy(a9.data(), aa, 9 * sizeof(double));
    w.rec(a9);
    w.text("atom: Nelement Natom / aa_atom / ki_atom /zn_atom");
    const int32_t na[2] = {nelement, natom};
    w.rec(na, 8);
    w.rec(aa_atom, sizeof(double) * 3 * natom);
    w.rec(ki_atom, sizeof(int32_t) * natom);
    w.rec(zn_atom, sizeof(int32_t) * nelement);
    w.text("Fermi energy");
    w.rec(&efermi, 8);
    for (int s = 0; s < MSP; ++s) {
        w.rec(rho + (size_t)s * ML0, 8 * (size_t)ML0);
        w.rec(vloc + (size_t)s * ML0, 8 * (size_t)ML0);
        if (s == 0) w.rec(vh, 8 * (size_t)ML0);
        w.rec(vxc + (size_t)s * ML0, 8 * (size_t)ML0);
    }
    if (w.close()) return set_error(kErrArg, "write_vrho: write to %s failed", file);
    return kOK;
}

int rsdft_gpu_read_vrho(const char *file, int N1, int N2, int N3, int ML0, int MSP, double *rho,
                        double *vloc, double *vh, double *vxc, double *efermi)
{
    const Grid3 g{N1, N2, N3};
    RS_TRY(check_grid(N1, N2, N3, ML0));
    if (!file || MSP < 1 || MSP > 2) return set_error(kErrArg, "read_vrho: bad argument");
    RecordReader r(file);
    if (!r.open()) return set_error(kErrArg, "read_vrho: cannot open %s", file);
    std::vector<unsigned char> b;
    if (!r.next(b)) return set_error(kErrArg, "read_vrho: %s: empty", file);
    const bool versioned = b.size() >= 7 && std::memcmp(b.data(), "version", 7) == 0;
    if (!versioned) r.rewind();
    int32_t head[4];
    if (!r.next(b) || !take(b, 4, head)) return set_error(kErrArg, "read_vrho: bad header");
    if (head[0] != ML0 || head[1] != N1 || head[2] != N2 || head[3] != N3)
        return set_error(kErrArg, "read_vrho: file grid %d (%dx%dx%d), this grid %d (%dx%dx%d)",
                         head[0], head[1], head[2], head[3], ML0, N1, N2, N3);
    std::vector<int32_t> LL(3 * (size_t)ML0);
    if (!r.next(b) || !take(b, LL.size(), LL.data()))
        return set_error(kErrArg, "read_vrho: bad grid map");
    std::vector<long> pos;
    RS_TRY(file_order(LL, g, pos));
    double ef = 0.0;
    if (versioned) {
        for (int k = 0; k < 8; ++k)
            if (!r.next(b)) return set_error(kErrArg, "read_vrho: bad header");
        if (!r.next(b) || !take(b, 1, &ef)) return set_error(kErrArg, "read_vrho: bad efermi");
    }
    if (efermi) *efermi = ef;
    auto field = [&](double *dst, const char *what) -> int {
        if (!r.next(b) || b.size() != 8 * (size_t)ML0)
            return set_error(kErrArg, "read_vrho: %s record missing or of the wrong size", what);
        const double *v = reinterpret_cast<const double *>(b.data());
        if (dst)
            for (long i = 0; i < ML0; ++i) dst[pos[i]] = v[i];
        return kOK;
    };
    for (int s = 0; s < MSP; ++s) {
        RS_TRY(field(rho ? rho + (size_t)s * ML0 : nullptr, "rho"));
        RS_TRY(field(vloc ? vloc + (size_t)s * ML0 : nullptr, "Vloc"));
        if (s == 0) RS_TRY(field(vh, "Vh"));
        RS_TRY(field(vxc ? vxc + (size_t)s * ML0 : nullptr, "Vxc"));
    }
    return kOK;
}

}  // extern "C"
