// fc_restart.hip -- the restart file of one flow rank (include/fc_restart.h):
// write_restart / read_restart (MAIN.F90:983-1122) over host or
// device-resident arrays.
//
// The records are Fortran unformatted sequential records as the reference's
// compiler lays them out: a 4-byte little-endian length before and after the
// data, and a logical record longer than the 4-byte markers allow split into
// subrecords whose leading marker is negative when another one follows and
// whose trailing marker is negative when one precedes.  Device arrays are
// staged through two pinned host buffers: the device-to-host copy of one
// chunk runs while the previous chunk is written (and on reading, the file
// read of one chunk runs while the previous one is copied to the device).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fc_restart.h"
#include "fc_handle.h"

namespace {

long long g_max_sub = 2147483639LL;
constexpr size_t kChunk = size_t(32) << 20;

struct Span {
    void *p;
    size_t bytes;
    bool dev;
};

int io_error(const std::string &what) { return fc::set_error(-7, "fc_restart: " + what); }
int fmt_error(const std::string &what) { return fc::set_error(-4, "fc_restart: " + what); }
int hip_error(hipError_t e, const char *what)
{
    return fc::set_error(-3, std::string("fc_restart: ") + what + ": " + hipGetErrorString(e));
}

#define RCHK(x)                                          \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return hip_error(e_, #x);  \
    } while (0)

// pinned double buffer and a stream, only for device arrays
struct Staging {
    char *buf[2] = {nullptr, nullptr};
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int init()
    {
        for (int q = 0; q < 2; ++q) {
            RCHK(hipHostMalloc((void **)&buf[q], kChunk, hipHostMallocDefault));
            RCHK(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming));
        }
        RCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return 0;
    }
    ~Staging()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (int q = 0; q < 2; ++q) {
            if (buf[q]) (void)hipHostFree(buf[q]);
            if (ev[q]) (void)hipEventDestroy(ev[q]);
        }
        if (st) (void)hipStreamDestroy(st);
    }
};

// one logical record being written: markers inserted at subrecord ends
struct RecordWriter {
    FILE *fp;
    size_t total, pos = 0, sub_start = 0, sub_end = 0;
    long long sub = 0;
    int marker(long long v)
    {
        const int32_t m = (int32_t)v;
        return fwrite(&m, 4, 1, fp) == 1 ? 0 : io_error("write failed");
    }
    size_t next_len() const
    {
        const size_t left = total - pos;
        return left < (size_t)g_max_sub ? left : (size_t)g_max_sub;
    }
    int begin()
    {
        const size_t len = next_len();
        sub_start = 0;
        sub_end = len;
        return marker(len < total ? -(long long)len : (long long)len);
    }
    int put(const void *src, size_t n)
    {
        const char *s = static_cast<const char *>(src);
        while (n > 0) {
            if (pos == sub_end) {
                const long long len = (long long)(sub_end - sub_start);
                if (int rc = marker(sub > 0 ? -len : len)) return rc;
                ++sub;
                const size_t nl = next_len();
                sub_start = pos;
                sub_end = pos + nl;
                if (int rc = marker(pos + nl < total ? -(long long)nl : (long long)nl)) return rc;
            }
            const size_t k = (n < sub_end - pos) ? n : sub_end - pos;
            if (fwrite(s, 1, k, fp) != k) return io_error("write failed");
            pos += k;
            s += k;
            n -= k;
        }
        return 0;
    }
    int end()
    {
        if (pos != total) return fmt_error("internal record length mismatch");
        const long long len = (long long)(sub_end - sub_start);
        return marker(sub > 0 ? -len : len);
    }
};

int write_record(FILE *fp, const std::vector<Span> &items, Staging *stg)
{
    RecordWriter w{fp, 0};
    for (const Span &s : items) w.total += s.bytes;
    if (int rc = w.begin()) return rc;
    for (const Span &s : items) {
        if (!s.dev) {
            if (int rc = w.put(s.p, s.bytes)) return rc;
            continue;
        }
        // chunk c copies into buf[c % 2] while chunk c - 1 is written
        const char *src = static_cast<const char *>(s.p);
        const size_t nchunk = (s.bytes + kChunk - 1) / kChunk;
        size_t prev_len = 0;
        for (size_t c = 0; c <= nchunk; ++c) {
            size_t len = 0;
            if (c < nchunk) {
                len = (c + 1 < nchunk) ? kChunk : s.bytes - c * kChunk;
                RCHK(hipMemcpyAsync(stg->buf[c % 2], src + c * kChunk, len, hipMemcpyDeviceToHost, stg->st));
                RCHK(hipEventRecord(stg->ev[c % 2], stg->st));
            }
            if (c > 0) {
                RCHK(hipEventSynchronize(stg->ev[(c - 1) % 2]));
                if (int rc = w.put(stg->buf[(c - 1) % 2], prev_len)) return rc;
            }
            prev_len = len;
        }
    }
    return w.end();
}

// one logical record being read: the sizes must add up to the record's
struct RecordReader {
    FILE *fp;
    size_t pos = 0, sub_end = 0, sub_len = 0;
    long long sub = 0;
    bool more = false;
    int marker(int32_t *m) { return fread(m, 4, 1, fp) == 1 ? 0 : io_error("read failed (end of file)"); }
    int lead()
    {
        int32_t m;
        if (int rc = marker(&m)) return rc;
        more = m < 0;
        sub_len = (size_t)(m < 0 ? -(long long)m : (long long)m);
        sub_end = pos + sub_len;
        return 0;
    }
    int trail()
    {
        int32_t t;
        if (int rc = marker(&t)) return rc;
        const long long a = t < 0 ? -(long long)t : (long long)t;
        if ((size_t)a != sub_len || (t < 0) != (sub > 0)) return fmt_error("record markers do not match");
        return 0;
    }
    int begin() { return lead(); }
    int get(void *dst, size_t n)
    {
        char *d = static_cast<char *>(dst);
        while (n > 0) {
            if (pos == sub_end) {
                if (!more) return fmt_error("record shorter than the arrays of these dimensions");
                if (int rc = trail()) return rc;
                ++sub;
                if (int rc = lead()) return rc;
                continue;
            }
            const size_t k = (n < sub_end - pos) ? n : sub_end - pos;
            if (fread(d, 1, k, fp) != k) return io_error("read failed (end of file)");
            pos += k;
            d += k;
            n -= k;
        }
        return 0;
    }
    int end()
    {
        if (pos != sub_end || more) return fmt_error("record longer than the arrays of these dimensions");
        return trail();
    }
};

int read_record(FILE *fp, const std::vector<Span> &items, Staging *stg)
{
    RecordReader r{fp};
    if (int rc = r.begin()) return rc;
    for (const Span &s : items) {
        if (!s.dev) {
            if (int rc = r.get(s.p, s.bytes)) return rc;
            continue;
        }
        // chunk c is read into buf[c % 2] while chunk c - 1 is copied up
        char *dst = static_cast<char *>(s.p);
        const size_t nchunk = (s.bytes + kChunk - 1) / kChunk;
        for (size_t c = 0; c < nchunk; ++c) {
            const size_t len = (c + 1 < nchunk) ? kChunk : s.bytes - c * kChunk;
            if (c >= 2) RCHK(hipEventSynchronize(stg->ev[c % 2]));
            if (int rc = r.get(stg->buf[c % 2], len)) return rc;
            RCHK(hipMemcpyAsync(dst + c * kChunk, stg->buf[c % 2], len, hipMemcpyHostToDevice, stg->st));
            RCHK(hipEventRecord(stg->ev[c % 2], stg->st));
        }
        RCHK(hipStreamSynchronize(stg->st));
    }
    return r.end();
}

int flow_spans(int nx, int ny, int nz, const fc_restart_flow *f, std::vector<Span> &out)
{
    double *const d[26] = {f->u, f->v, f->w, f->p, f->face_u, f->face_v, f->face_w, f->pPrime,
                           f->bcxu, f->bcxv, f->bcxw, f->bcxp, f->bcyu, f->bcyv, f->bcyw, f->bcyp,
                           f->bczu, f->bczv, f->bczw, f->bczp, f->pgradx1, f->pgradx2, f->pgrady1,
                           f->pgrady2, f->pgradz1, f->pgradz2};
    for (double *p : d)
        if (!p) return fc::set_error(-1, "fc_restart: a flow array is NULL");
    if (!f->iblank || !f->ghostCellMark) return fc::set_error(-1, "fc_restart: iblank or ghostCellMark is NULL");
    const size_t X = (size_t)nx + 2, Y = (size_t)ny + 2, Z = (size_t)nz + 2;
    const size_t n3 = X * Y * Z;
    const size_t cnt[26] = {n3, n3, n3, n3, n3, n3, n3, n3,
                            2 * Y * Z, 2 * Y * Z, 2 * Y * Z, 2 * Y * Z,
                            X * 2 * Z, X * 2 * Z, X * 2 * Z, X * 2 * Z,
                            X * Y * 2, X * Y * 2, X * Y * 2, X * Y * 2,
                            Y * Z, Y * Z, X * Z, X * Z, X * Y, X * Y};
    const bool dev = f->on_device != 0;
    for (int a = 0; a < 26; ++a) out.push_back(Span{d[a], cnt[a] * sizeof(double), dev});
    out.push_back(Span{f->iblank, n3, dev});
    out.push_back(Span{f->ghostCellMark, n3, dev});
    return 0;
}

int body_spans(const fc_restart_body &b, std::vector<Span> &out)
{
    if (b.nPts < 1 || b.nElem < 1) return fc::set_error(-1, "fc_restart: a body needs nPts >= 1 and nElem >= 1");
    double *const d[9] = {b.xBodyMarker, b.yBodyMarker, b.zBodyMarker, b.uBodyMarker, b.vBodyMarker,
                          b.wBodyMarker, b.xMarkerForce, b.yMarkerForce, b.zMarkerForce};
    for (double *p : d) {
        if (!p) return fc::set_error(-1, "fc_restart: a body array is NULL");
        out.push_back(Span{p, sizeof(double) * (size_t)b.nPts, false});
    }
    if (!b.triElemNeig) return fc::set_error(-1, "fc_restart: triElemNeig is NULL");
    out.push_back(Span{b.triElemNeig, sizeof(int) * 3 * (size_t)b.nElem, false});
    // xcent ycent zcent angx angy angz vxcent vycent vzcent angvx angvy angvz:
    // four 3-vectors of the struct, one after the other
    double *const s[4] = {const_cast<double *>(b.cent), const_cast<double *>(b.ang), const_cast<double *>(b.vcent),
                          const_cast<double *>(b.angv)};
    for (double *p : s) out.push_back(Span{p, 3 * sizeof(double), false});
    return 0;
}

struct File {
    FILE *fp = nullptr;
    ~File()
    {
        if (fp) fclose(fp);
    }
};

}  // namespace

extern "C" int fc_restart_name(int lProc, char *out, int out_len)
{
    if (!out || out_len < 17 || lProc < 0 || lProc > 9999)
        return fc::set_error(-1, "fc_restart_name: lProc must be 0..9999 and out at least 17 bytes");
    std::snprintf(out, (size_t)out_len, "restart_%04d.dat", lProc);
    return 0;
}

extern "C" long long fc_restart_set_max_subrecord(long long bytes)
{
    const long long old = g_max_sub;
    if (bytes >= 8 && bytes <= 2147483639LL) g_max_sub = bytes;
    return old;
}

extern "C" int fc_restart_write(const char *path, int nx, int ny, int nz, double time, int ntime,
                                const fc_restart_flow *f, int nBody, const fc_restart_body *bodies)
{
    if (!path || !f || nx < 1 || ny < 1 || nz < 1 || nBody < 0 || (nBody > 0 && !bodies))
        return fc::set_error(-1, "fc_restart_write: bad argument");
    std::vector<Span> flow;
    if (int rc = flow_spans(nx, ny, nz, f, flow)) return rc;
    std::vector<std::vector<Span>> bs((size_t)nBody);
    for (int b = 0; b < nBody; ++b)
        if (int rc = body_spans(bodies[b], bs[(size_t)b])) return rc;
    Staging stg;
    if (f->on_device)
        if (int rc = stg.init()) return rc;
    File file;
    file.fp = std::fopen(path, "wb");
    if (!file.fp) return io_error(std::string("cannot open ") + path + " for writing");
    int32_t nt = ntime;
    const std::vector<Span> head = {Span{&time, 8, false}, Span{&nt, 4, false}};
    if (int rc = write_record(file.fp, head, nullptr)) return rc;
    if (int rc = write_record(file.fp, flow, &stg)) return rc;
    for (int b = 0; b < nBody; ++b)
        if (int rc = write_record(file.fp, bs[(size_t)b], nullptr)) return rc;
    if (std::fflush(file.fp) != 0) return io_error("flush failed");
    return 0;
}

extern "C" int fc_restart_read(const char *path, int nx, int ny, int nz, double *time, int *ntime_start,
                               const fc_restart_flow *f, int nBody, fc_restart_body *bodies)
{
    if (!path || !f || !time || !ntime_start || nx < 1 || ny < 1 || nz < 1 || nBody < 0 || (nBody > 0 && !bodies))
        return fc::set_error(-1, "fc_restart_read: bad argument");
    std::vector<Span> flow;
    if (int rc = flow_spans(nx, ny, nz, f, flow)) return rc;
    std::vector<std::vector<Span>> bs((size_t)nBody);
    for (int b = 0; b < nBody; ++b)
        if (int rc = body_spans(bodies[b], bs[(size_t)b])) return rc;
    Staging stg;
    if (f->on_device)
        if (int rc = stg.init()) return rc;
    File file;
    file.fp = std::fopen(path, "rb");
    if (!file.fp) return io_error(std::string("cannot open ") + path + " for reading");
    double t = 0.0;
    int32_t nt = 0;
    const std::vector<Span> head = {Span{&t, 8, false}, Span{&nt, 4, false}};
    if (int rc = read_record(file.fp, head, nullptr)) return rc;
    if (int rc = read_record(file.fp, flow, &stg)) return rc;
    for (int b = 0; b < nBody; ++b)
        if (int rc = read_record(file.fp, bs[(size_t)b], nullptr)) return rc;
    *time = t;
    *ntime_start = nt;
    return 0;
}
