// fc_poisson.hip -- host orchestration and C-ABI (include/fc_poisson.h) of
// the gfx950 pressure-Poisson hot path: hierarchy construction (MG_initial),
// level masks (mg_prepare), GCM list preparation, and the solve_poisson /
// mg_solve / GCM_ghostcell_pressure control flow driving the kernels of
// fc_kernels.hip on one HIP stream per handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fc_poisson.h"
#include "fc_handle.h"
#include "fc_kernels.h"

using namespace fc;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return fail(-3, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

// Scratch of the allocation-free entry point fc_pson_rhs_noalloc: the two
// per-row partial sums of rhs_poisson's reductions for up to
// FC_NOALLOC_MAX_ROWS interior (j,k) rows (4 MiB).  Every GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// is only a stand-in for the symbol: its device address is taken per device
// with hipGetSymbolAddress (noalloc_scratch) and only that address is used.
static __device__ __align__(256) double scratch[2 * FC_NOALLOC_MAX_ROWS];

namespace {

// device address of scratch on each device, resolved once per device while
// that device is current (at handle creation)
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int resolve_scratch(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(-1, "device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return 0;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(-3, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)ptr;
    return 0;
}

double *scratch_of(int dev)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[dev];
}

const int MAXLEV = 10;

struct HostMet {
    std::vector<double> x, y, z, xc, yc, zc, dxinv, dyinv, dzinv, dxcinv, dycinv, dzcinv;
};

// make_grid / metrics (CORE_GRID.F90:106-150)
void grid_metrics(int n, const std::vector<double> &x, std::vector<double> &xc, std::vector<double> &dx,
                  std::vector<double> &dxinv, std::vector<double> &dxc, std::vector<double> &dxcinv)
{
    xc.assign(n + 2, 0.0); dx.assign(n + 2, 0.0); dxinv.assign(n + 2, 0.0);
    dxc.assign(n + 2, 0.0); dxcinv.assign(n + 2, 0.0);
    for (int i = 0; i <= n; ++i) {
        xc[i] = 0.5 * (x[i + 1] + x[i]);
        dx[i] = x[i + 1] - x[i];
        dxinv[i] = 1.0 / dx[i];
    }
    xc[n + 1] = x[n + 1];
    for (int i = 1; i <= n; ++i) {
        dxc[i] = xc[i] - xc[i - 1];
        dxcinv[i] = 1.0 / dxc[i];
    }
}

// MG_initial level count (CORE_MG.F90:69-102)
int auto_levels(int n, int mg)
{
    if (mg == 0 || mg > MAXLEV) {
        int ii = n - 1, N;
        for (N = 1; N <= MAXLEV; ++N) {
            ii = ii / 2;
            if (ii >= 4) continue;
            break;
        }
        return std::min(N, MAXLEV);
    }
    return mg;
}

int coarsen(int n) { return (n % 2 == 0) ? n / 2 + 1 : (n + 1) / 2; }

// coarse metrics of one axis (CORE_MG.F90:234-330); ghost nodes mirrored
void coarse_axis(int nc, int nf, const std::vector<double> &xf, double x1, double xn, std::vector<double> &x,
                 std::vector<double> &xc, std::vector<double> &dxinv, std::vector<double> &dxcinv)
{
    x.assign(nc + 2, 0.0); xc.assign(nc + 2, 0.0); dxinv.assign(nc + 2, 0.0); dxcinv.assign(nc + 2, 0.0);
    for (int i = 1; i <= nc - 1; ++i) x[i] = xf[(nc == nf) ? i : 2 * i - 1];
    x[nc] = xf[nf];
    x[0] = 2.0 * x[1] - x[2];
    x[nc + 1] = 2.0 * x[nc] - x[nc - 1];
    for (int i = 1; i <= nc - 1; ++i) {
        dxinv[i] = 1.0 / (x[i + 1] - x[i]);
        xc[i] = 0.5 * (x[i] + x[i + 1]);
    }
    dxinv[0] = 1.0 / (x[1] - x[0]);
    dxinv[nc] = 1.0 / (x[nc + 1] - x[nc]);
    xc[0] = 2.0 * x1 - xc[1];
    xc[nc] = 2.0 * xn - xc[nc - 1];
    xc[nc + 1] = x[nc + 1];
    for (int i = 1; i <= nc; ++i) dxcinv[i] = 1.0 / (xc[i] - xc[i - 1]);
}

template <typename T>
struct DevMet {
    T *buf = nullptr;
    Metrics<T> m{};
};

long pad16(long n) { return (n + 15) / 16 * 16; }

}  // namespace

struct fc_pson {
    fc_pson_params prm{};
    int device = 0;              // the device the handle was created on
    int nx = 0, ny = 0, nz = 0;
    hipStream_t st = nullptr;
    int nlev = 0;
    std::vector<LevelDims> dims;
    std::vector<HostMet> hm;
    std::vector<double> dx1, dy1, dz1, dxc1, dyc1, dzc1;    // level-1 widths
    std::vector<DevMet<double>> m64;
    std::vector<DevMet<float>> m32;
    std::vector<LevelPtrs<double>> L64;
    std::vector<LevelPtrs<float>> L32;
    std::vector<int8_t *> ibk, igk;
    // per coarse level l >= 1: rflag = some restriction child (level l-1) is
    // ibk; pflag(i1,j1,k1) = some cell of the prolongation block starting at
    // (i1,j1,k1) on level l is ibk (both static per mask set)
    std::vector<int8_t *> rflag, pflag;
    // correction V-cycles with bodies as walls: per level the cell mask byte
    // (bit 0 held, bits 1..6 wall faces) and, per coarse level, the flag of
    // interpolation blocks that hold a held coarse cell
    std::vector<int8_t *> wmask, wpflag;
    double *p = nullptr, *div = nullptr, *pprev = nullptr, *ptemp = nullptr;
    double *d_l1w = nullptr;    // device dx,dy,dz, dxinv,dyinv,dzinv, dxcinv,... of level 1
    const double *l1met[6] = {};   // dxcinv, dycinv, dzcinv, dxinv, dyinv, dzinv
    const double *l1rhs[6] = {};   // dxinv, dyinv, dzinv, dx, dy, dz
    double *pg[6] = {};
    double outflow_area = 1.0;
    double outflow_corr = 0.0;    // velocity correction of the last rhs_poisson (zero-gradient faces)
    // host copies of the level-1 masks (Fortran layout)
    std::vector<int8_t> h_iblank, h_ghost, h_dead;
    bool have_masks = false;
    // GCM
    struct {
        std::vector<int> iG, jG, kG, iC, jC, kC, iD, jD, kD, iDC, jDC, kDC;
        std::vector<double> cN, alpha, dpdnG, cDN, plD, dpdnD, ipD;
    } hg;
    bool gcm_dirty = true;
    GcmDev g{};
    void *gbuf = nullptr;
    long *off8 = nullptr;
    int *gijk = nullptr;
    double *gvalD = nullptr, *gvalG = nullptr, *reslap = nullptr;
    double *gdelta = nullptr;    // change of each relation's target in the last exact pass
    // topological schedule of the GCM relations (mix_GC_form = 0)
    GcmNode *gnode = nullptr;    // the relations as records in schedule order
    std::vector<int> glevel_off;
    bool gsched_ok = false;
    int gsched_levels = 0;
    // scratch
    double *partial = nullptr;
    long npartial = 0;
    double *scal = nullptr;      // device scalars
    double *hscal = nullptr;     // pinned host mirror
    size_t dev_bytes = 0;
    bool mixed_ready = false;
    // FP64 correction hierarchy of the FP64 fold (level 1: its own e and rhs;
    // coarse levels shared with L64), built on first use
    std::vector<LevelPtrs<double>> L64c;
    // Anderson acceleration state (device resident)
    int aa_m = 0;
    double *aa_x = nullptr, *aa_g = nullptr;
    float *aa_fprev = nullptr, *aa_c = nullptr;
    AaCols aa_cols{};
    int *aa_ring = nullptr;
    // device memory pool (handles of a request server) and stream ownership
    Arena *arena = nullptr;
    bool own_stream = true;
    int *d_glevel = nullptr;     // GCM topological level offsets (batched kernels)
    // finest-level sweep timing (events on the handle's stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;         // pairs: start, stop
    size_t ev_used = 0;
    double sweep_ms = 0.0;
    long sweep_launches = 0;
    // velocity projection (allocated at its first use): u, v, w, face_u,
    // face_v, face_w as six padded level-1 fields; bcxu/bcyv/bczw in the
    // host layout; the metrics; the update_pressure zero mask; the ghost list
    double *vel = nullptr, *velbc = nullptr, *projmet = nullptr;
    ProjMet pm{};
    int8_t *pzero = nullptr;
    bool pzero_dirty = true;
    GhostVel *gvel = nullptr;
    int ngvel = 0;
    bool gvel_dirty = true;
    double proj_ms[4] = {};       // kernel times of the last timed projection
    // solve_ad's outer velocity conditions (set_outer_velocity_bc): face
    // values and frequencies (ux1, ux2, vx1, ..., wz2), the channel-profile
    // bounds, and level 1's node y on the device
    double vbc_val[18] = {}, vbc_freq[18] = {}, vbc_y0 = 0.0, vbc_y1 = 1.0;
    double *d_ynode = nullptr;
    // GCM_ghostcell_velocity: coeffGCMD | coeffGCMDeadD | u_BI | u_BI dead,
    // probeLengthNormalizedD, and 3 max(nG, nD) doubles of step values
    double *gv_buf = nullptr, *gv_vals = nullptr;
    double gv_probe = 2.0;
    bool have_gcm_vel = false;
    // FP64 direct solve of the coarsest level of the wall correction cycles:
    // cd_M (n x n) maps the level's rhs at the free cells cd_q to e there;
    // cd_state 0 = not built, 1 = in use, -1 = not applicable
    double *cd_M = nullptr;
    long *cd_q = nullptr;
    int cd_n = 0, cd_state = 0;
    // advection-diffusion: set_solve_ad coefficients, the right-hand sides
    // nlu, nlv, nlw (three padded fields), reduction scratch
    double *ad_coefs = nullptr, *ad_rhs = nullptr, *ad_partial = nullptr;
    AdMet am{};
    bool ad_ready = false;
    double ad_ms = 0.0;           // time of the last timed itsolv_ad or rhs_advec_diff call
    double ad_reinv = 0.0;        // nu of set_solve_ad
    // rhs_advec_diff: upwind weights per axis (x | y | z, index 0..n+1) and
    // viscTot as a padded level-1 field (unused while have_visc is false:
    // viscTot = reinv, MAIN.F90:926)
    UpwW *upw = nullptr;
    double *visc_tot = nullptr;
    bool have_visc = false;
};

namespace {

template <typename T>
int dalloc(fc_pson *h, T **p, size_t n)
{
    size_t b = sizeof(T) * std::max<size_t>(n, 1);
    if (h->arena) {
        *p = (T *)h->arena->alloc(b);
        if (!*p) return fail(-6, "device memory pool exhausted");
    } else {
        hipError_t e = hipMalloc((void **)p, b);
        if (e != hipSuccess) return fail(-4, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    h->dev_bytes += b;
    hipMemsetAsync(*p, 0, b, h->st);
    return 0;
}

void dfree(fc_pson *h, void *p)
{
    if (!p) return;
    if (h->arena) h->arena->free(p);
    else hipFree(p);
}

template <typename T>
int upload_metrics(fc_pson *h, int lev, DevMet<T> &dm)
{
    const LevelDims &d = h->dims[lev];
    const HostMet &hm = h->hm[lev];
    int nx2 = d.nx + 2, ny2 = d.ny + 2, nz2 = d.nz + 2;
    // each array starts on a 4-element boundary and has at least 4 zero
    // elements past index n+1, so a kernel may load 4 consecutive entries
    // from any multiple of 4 up to n+1 with one aligned vector load
    auto span = [](int n) { return (size_t)((n + 3) / 4 * 4 + 4); };
    size_t tot = 7 * (span(nx2) + span(ny2) + span(nz2));
    std::vector<T> v(tot, (T)0);
    size_t o = 0;
    auto put = [&](const std::vector<double> &a, int n) {
        size_t start = o;
        for (int i = 0; i < n; ++i) v[o + i] = (T)a[i];
        o += span(n);
        return start;
    };
    size_t ox = put(hm.x, nx2), oy = put(hm.y, ny2), oz = put(hm.z, nz2);
    size_t oxc = put(hm.xc, nx2), oyc = put(hm.yc, ny2), ozc = put(hm.zc, nz2);
    size_t odx = put(hm.dxinv, nx2), ody = put(hm.dyinv, ny2), odz = put(hm.dzinv, nz2);
    size_t ocx = put(hm.dxcinv, nx2), ocy = put(hm.dycinv, ny2), ocz = put(hm.dzcinv, nz2);
    auto coefs = [&](const std::vector<double> &dcinv, const std::vector<double> &dinv, int n, size_t &om,
                     size_t &op, size_t &oc) {
        std::vector<double> am(n, 0.0), ap(n, 0.0), ac(n, 0.0);
        for (int i = 0; i + 1 < n; ++i) {
            am[i] = dcinv[i] * dinv[i];
            ap[i] = dcinv[i + 1] * dinv[i];
            ac[i] = -(am[i] + ap[i]);
        }
        om = put(am, n); op = put(ap, n); oc = put(ac, n);
    };
    size_t oamx, oapx, oacx, oamy, oapy, oacy, oamz, oapz, oacz;
    coefs(hm.dxcinv, hm.dxinv, nx2, oamx, oapx, oacx);
    coefs(hm.dycinv, hm.dyinv, ny2, oamy, oapy, oacy);
    coefs(hm.dzcinv, hm.dzinv, nz2, oamz, oapz, oacz);
    int rc = dalloc(h, &dm.buf, tot);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(dm.buf, v.data(), sizeof(T) * tot, hipMemcpyHostToDevice, h->st));
    T *b = dm.buf;
    dm.m = Metrics<T>{b + ox, b + oy, b + oz, b + oxc, b + oyc, b + ozc, b + odx, b + ody, b + odz,
                      b + ocx, b + ocy, b + ocz, b + oamx, b + oapx, b + oacx, b + oamy, b + oapy, b + oacy,
                      b + oamz, b + oapz, b + oacz};
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

BcInfo bcinfo(const fc_pson *h, int lev, bool homo)
{
    BcInfo b;
    for (int f = 0; f < 6; ++f) { b.pbc[f] = h->prm.pbc[f]; b.ppp[f] = h->prm.ppp[f]; }
    b.pgx1 = h->pg[0]; b.pgx2 = h->pg[1]; b.pgy1 = h->pg[2]; b.pgy2 = h->pg[3]; b.pgz1 = h->pg[4]; b.pgz2 = h->pg[5];
    b.dxc1 = h->dxc1[1]; b.dxcn = h->dxc1[h->nx];
    b.dyc1 = h->dyc1[1]; b.dycn = h->dyc1[h->ny];
    b.dzc1 = h->dzc1[1]; b.dzcn = h->dzc1[h->nz];
    b.homo = (homo || lev > 0) ? 1 : 0;
    b.fnx = h->nx; b.fny = h->ny;
    const int *v = h->prm.bc_vel;
    b.per = ((v[0] == FC_BC_TYPE_PERIODIC && v[1] == FC_BC_TYPE_PERIODIC) ? 1 : 0) |
            ((v[2] == FC_BC_TYPE_PERIODIC && v[3] == FC_BC_TYPE_PERIODIC) ? 2 : 0);
    return b;
}

// Fortran layout (nx+2 contiguous) <-> padded device layout.  The rows are
// re-pitched on the host and cross PCIe as one contiguous copy: a pitched
// copy from pageable memory goes row by row, which at 10^4-10^5 rows costs
// far more than the bytes.  Both return after the copy is complete.
// f(k) for k = lo..hi on up to 16 host threads (whole k-planes per call)
// when the planes hold at least 8M cells together: the host-side mask
// preparation of every level at large grids (small grids, such as the
// serving requests built on worker threads, stay on the calling thread)
template <typename F>
void par_planes(int lo, int hi, long plane_cells, F &&f)
{
    const int n = hi - lo + 1;
    if (n <= 0) return;
    int T = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (T > n) T = n;
    if (T <= 1 || (long)n * plane_cells < (8L << 20)) {
        for (int k = lo; k <= hi; ++k) f(k);
        return;
    }
    std::atomic<int> next(lo);
    std::vector<std::thread> th;
    th.reserve((size_t)T);
    for (int t = 0; t < T; ++t)
        th.emplace_back([&]() {
            for (int k; (k = next.fetch_add(1)) <= hi;) f(k);
        });
    for (auto &x : th) x.join();
}

template <typename T>
int to_device(fc_pson *h, T *dst, const T *src, int lev = 0)
{
    const LevelDims &d = h->dims[lev];
    const size_t rows = (size_t)(d.ny + 2) * (d.nz + 2), w = (size_t)d.nx + 2;
    // staging in the padded layout: every row copied, its padding zeroed
    // (no separate fill of the whole buffer)
    std::unique_ptr<T[]> st(new T[(size_t)d.n]);
    const size_t rj = (size_t)d.ny + 2;
    const size_t pad = (size_t)d.px - w;
    par_planes(0, d.nz + 1, (long)rj * (long)w, [&](int k) {
        for (size_t r = (size_t)k * rj; r < (size_t)(k + 1) * rj; ++r) {
            memcpy(st.get() + r * d.px, src + r * w, sizeof(T) * w);
            if (pad) memset(st.get() + r * d.px + w, 0, sizeof(T) * pad);
        }
    });
    for (size_t q = rows * (size_t)d.px; q < (size_t)d.n; ++q) st[q] = (T)0;
    HIPCHK(hipMemcpyAsync(dst, st.get(), sizeof(T) * (size_t)d.n, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

template <typename T>
int to_host(fc_pson *h, T *dst, const T *src, int lev = 0)
{
    const LevelDims &d = h->dims[lev];
    const size_t rows = (size_t)(d.ny + 2) * (d.nz + 2), w = (size_t)d.nx + 2;
    std::vector<T> st((size_t)d.n);
    HIPCHK(hipMemcpyAsync(st.data(), src, sizeof(T) * (size_t)d.n, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t r = 0; r < rows; ++r) memcpy(dst + r * w, st.data() + r * d.px, sizeof(T) * w);
    return 0;
}

int ensure_mixed(fc_pson *h)
{
    if (h->mixed_ready) return 0;
    h->L32.resize(h->nlev);
    for (int l = 0; l < h->nlev; ++l) {
        int rc;
        if ((rc = dalloc(h, &h->L32[l].phi, h->dims[l].n))) return rc;
        if ((rc = dalloc(h, &h->L32[l].rhs, h->dims[l].n))) return rc;
        if ((rc = dalloc(h, &h->L32[l].res, h->dims[l].n))) return rc;
    }
    HIPCHK(hipStreamSynchronize(h->st));
    h->mixed_ready = true;
    return 0;
}

// host-side mg_prepare (CORE_MG.F90:1365-1555) and upload of the masks
int prepare_masks(fc_pson *h)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    std::vector<std::vector<int8_t>> IBK(h->nlev), IGK(h->nlev);
    auto F = [&](int l, int i, int j, int k) {
        const LevelDims &d = h->dims[l];
        return (long)i + (long)(d.nx + 2) * ((long)j + (long)(d.ny + 2) * k);
    };
    {
        long n = (long)(nx + 2) * (ny + 2) * (nz + 2);
        IBK[0].assign(n, 0);
        IGK[0].assign(n, 0);
        par_planes(0, nz + 1, (long)(nx + 2) * (ny + 2), [&](int k) {
            for (int j = 0; j <= ny + 1; ++j)
                for (int i = 0; i <= nx; ++i) {
                    long q = F(0, i, j, k);
                    int8_t ib = h->h_iblank[q], gh = h->h_ghost[q];
                    IBK[0][q] = ib;
                    IGK[0][q] = (ib == 1 || gh == 1) ? 1 : 0;
                }
        });
    }
    for (int l = 1; l < h->nlev; ++l) {
        const LevelDims &c = h->dims[l], &f = h->dims[l - 1];
        const HostMet &mf = h->hm[l - 1];
        long n = (long)(c.nx + 2) * (c.ny + 2) * (c.nz + 2);
        IBK[l].assign(n, 0);
        IGK[l].assign(n, 0);
        par_planes(1, c.nz - 1, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
            for (int j = 1; j <= c.ny - 1; ++j)
                for (int i = 1; i <= c.nx - 1; ++i) {
                    int i1 = (1 - c.ic) * i + c.ic * (2 * i - 1), i2 = (1 - c.ic) * i + c.ic * (2 * i);
                    int j1 = (1 - c.jc) * j + c.jc * (2 * j - 1), j2 = (1 - c.jc) * j + c.jc * (2 * j);
                    int k1 = (1 - c.kc) * k + c.kc * (2 * k - 1), k2 = (1 - c.kc) * k + c.kc * (2 * k);
                    double vol0 = 0.0, volibk = 0.0;
                    int8_t g = 0;
                    for (int kk = k1; kk <= k2; ++kk)
                        for (int jj = j1; jj <= j2; ++jj)
                            for (int ii = i1; ii <= i2; ++ii) {
                                long q = F(l - 1, ii, jj, kk);
                                if (IGK[l - 1][q] == 1) g = 1;
                                double v = (mf.x[ii + 1] - mf.x[ii]) * (mf.y[jj + 1] - mf.y[jj]) * (mf.z[kk + 1] - mf.z[kk]);
                                vol0 = vol0 + v;
                                if (IBK[l - 1][q] == 1) volibk = volibk + v;
                            }
                    long q = F(l, i, j, k);
                    IGK[l][q] = g;
                    if (volibk / vol0 >= 0.5) IBK[l][q] = 1;
                }
        });
        (void)f;
    }
    std::vector<std::vector<int8_t>> RF(h->nlev), PF(h->nlev);
    for (int l = 1; l < h->nlev; ++l) {
        const LevelDims &c = h->dims[l];
        long n = (long)(c.nx + 2) * (c.ny + 2) * (c.nz + 2);
        RF[l].assign(n, 0);
        PF[l].assign(n, 0);
        par_planes(1, c.nz - 1, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
            for (int j = 1; j <= c.ny - 1; ++j)
                for (int i = 1; i <= c.nx - 1; ++i) {
                    int i1 = (1 - c.ic) * i + c.ic * (2 * i - 1), i2 = (1 - c.ic) * i + c.ic * (2 * i);
                    int j1 = (1 - c.jc) * j + c.jc * (2 * j - 1), j2 = (1 - c.jc) * j + c.jc * (2 * j);
                    int k1 = (1 - c.kc) * k + c.kc * (2 * k - 1), k2 = (1 - c.kc) * k + c.kc * (2 * k);
                    int8_t a = 0;
                    for (int kk = k1; kk <= k2; ++kk)
                        for (int jj = j1; jj <= j2; ++jj)
                            for (int ii = i1; ii <= i2; ++ii) a |= (IBK[l - 1][F(l - 1, ii, jj, kk)] == 1);
                    RF[l][F(l, i, j, k)] = a;
                }
        });
        par_planes(0, c.nz, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
            for (int j = 0; j <= c.ny; ++j)
                for (int i = 0; i <= c.nx; ++i) {
                    int8_t a = 0;
                    for (int kk = k; kk <= std::min(k + c.kc, c.nz + 1); ++kk)
                        for (int jj = j; jj <= std::min(j + c.jc, c.ny + 1); ++jj)
                            for (int ii = i; ii <= std::min(i + c.ic, c.nx + 1); ++ii) a |= (IBK[l][F(l, ii, jj, kk)] == 1);
                    PF[l][F(l, i, j, k)] = a;
                }
        });
    }
    // Wall masks of the correction V-cycles.  A cell is free on level 1 if it
    // is a fluid cell (iblank = 0: flag and ghost cells); on a coarse level if
    // one of its restriction children is free.  A free cell's face is a wall
    // when the interior cell across it is held; faces on the domain boundary
    // keep the outer condition (the ghost layer reflection).  Cells outside
    // the interior are marked free with no walls (never read as rows).
    std::vector<std::vector<int8_t>> WM(h->nlev), WPF(h->nlev);
    {
        std::vector<std::vector<uint8_t>> fr(h->nlev);
        for (int l = 0; l < h->nlev; ++l) {
            const LevelDims &c = h->dims[l];
            long n = (long)(c.nx + 2) * (c.ny + 2) * (c.nz + 2);
            fr[l].assign(n, 1);
            par_planes(1, c.nz - 1, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
                for (int j = 1; j <= c.ny - 1; ++j)
                    for (int i = 1; i <= c.nx - 1; ++i) {
                        uint8_t f = 0;
                        if (l == 0) {
                            f = h->h_iblank[F(0, i, j, k)] == 0;
                        } else {
                            int i1 = (1 - c.ic) * i + c.ic * (2 * i - 1), i2 = (1 - c.ic) * i + c.ic * (2 * i);
                            int j1 = (1 - c.jc) * j + c.jc * (2 * j - 1), j2 = (1 - c.jc) * j + c.jc * (2 * j);
                            int k1 = (1 - c.kc) * k + c.kc * (2 * k - 1), k2 = (1 - c.kc) * k + c.kc * (2 * k);
                            int nf = 0, nc = 0;
                            for (int kk = k1; kk <= k2; ++kk)
                                for (int jj = j1; jj <= j2; ++jj)
                                    for (int ii = i1; ii <= i2; ++ii) { nf += fr[l - 1][F(l - 1, ii, jj, kk)]; nc++; }
                            const int rule = h->prm.vc_walls;
                            f = (rule == 3) ? (nf == nc) : (rule == 2) ? (2 * nf >= nc) : (nf > 0);
                        }
                        fr[l][F(l, i, j, k)] = f;
                    }
            });
            WM[l].assign(n, 0);
            auto held = [&](int i, int j, int k) {
                bool inner = i >= 1 && i <= c.nx - 1 && j >= 1 && j <= c.ny - 1 && k >= 1 && k <= c.nz - 1;
                return inner && !fr[l][F(l, i, j, k)];
            };
            par_planes(1, c.nz - 1, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
                for (int j = 1; j <= c.ny - 1; ++j)
                    for (int i = 1; i <= c.nx - 1; ++i) {
                        int8_t b = 0;
                        if (held(i, j, k)) b = 1;
                        else {
                            if (held(i - 1, j, k)) b |= 2;
                            if (held(i + 1, j, k)) b |= 4;
                            if (held(i, j - 1, k)) b |= 8;
                            if (held(i, j + 1, k)) b |= 16;
                            if (held(i, j, k - 1)) b |= 32;
                            if (held(i, j, k + 1)) b |= 64;
                        }
                        WM[l][F(l, i, j, k)] = b;
                    }
            });
            if (l > 0) {
                WPF[l].assign(n, 0);
                par_planes(0, c.nz, (long)(c.nx + 2) * (c.ny + 2), [&](int k) {
                    for (int j = 0; j <= c.ny; ++j)
                        for (int i = 0; i <= c.nx; ++i) {
                            int8_t a = 0;
                            for (int kk = k; kk <= std::min(k + c.kc, c.nz + 1); ++kk)
                                for (int jj = j; jj <= std::min(j + c.jc, c.ny + 1); ++jj)
                                    for (int ii = i; ii <= std::min(i + c.ic, c.nx + 1); ++ii)
                                        a |= (WM[l][F(l, ii, jj, kk)] & 1);
                            WPF[l][F(l, i, j, k)] = a;
                        }
                });
            }
        }
    }
    for (int l = 0; l < h->nlev; ++l) {
        int rc;
        if ((rc = to_device(h, h->wmask[l], WM[l].data(), l))) return rc;
        if (l > 0 && (rc = to_device(h, h->wpflag[l], WPF[l].data(), l))) return rc;
        if ((rc = to_device(h, h->ibk[l], IBK[l].data(), l))) return rc;
        if ((rc = to_device(h, h->igk[l], IGK[l].data(), l))) return rc;
        if (l > 0) {
            if ((rc = to_device(h, h->rflag[l], RF[l].data(), l))) return rc;
            if ((rc = to_device(h, h->pflag[l], PF[l].data(), l))) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

// GCM lists -> device (linear offsets on the padded level-1 layout, self
// rows, inverse-distance branch decisions of GCM_GHOSTCELL_UPDATE.F90:
// 459-506, which depend on geometry only)
int finalize_gcm(fc_pson *h)
{
    if (!h->gcm_dirty) return 0;
    if (!h->have_masks) return fail(-1, "fc_pson_set_masks must be called before solving");
    const LevelDims &d = h->dims[0];
    auto lin = [&](int i, int j, int k) { return (long)i + d.sy * j + d.sz * k; };
    auto F = [&](int i, int j, int k) { return (long)i + (long)(h->nx + 2) * ((long)j + (long)(h->ny + 2) * k); };
    static const int incI[8] = {0, 0, 1, 1, 0, 0, 1, 1}, incJ[8] = {0, 1, 1, 0, 0, 1, 1, 0},
                     incK[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    auto &hg = h->hg;
    int nG = (int)hg.iG.size(), nD = (int)hg.iD.size();
    auto inb = [&](int i, int j, int k) { return i >= 0 && i <= h->nx && j >= 0 && j <= h->ny && k >= 0 && k <= h->nz; };
    std::vector<long> gT(nG), gB(nG), dT(nD), dB(nD);
    std::vector<int> gS(nG), dS(nD), dA(nD), dI(nD), gijk(3 * (size_t)std::max(nG, 1));
    std::vector<double> dW(8 * (size_t)std::max(nD, 1), 0.0);
    for (int n = 0; n < nG; ++n) {
        if (!inb(hg.iG[n], hg.jG[n], hg.kG[n]) || !inb(hg.iC[n], hg.jC[n], hg.kC[n]) ||
            !inb(hg.iC[n] + 1, hg.jC[n] + 1, hg.kC[n] + 1))
            return fail(-1, "ghost-cell or host-cell index out of range at n=" + std::to_string(n + 1));
        gT[n] = lin(hg.iG[n], hg.jG[n], hg.kG[n]);
        gB[n] = lin(hg.iC[n], hg.jC[n], hg.kC[n]);
        gS[n] = -1;
        for (int r = 0; r < 8; ++r)
            if (hg.iC[n] + incI[r] == hg.iG[n] && hg.jC[n] + incJ[r] == hg.jG[n] && hg.kC[n] + incK[r] == hg.kG[n]) gS[n] = r;
        gijk[3 * n] = hg.iG[n]; gijk[3 * n + 1] = hg.jG[n]; gijk[3 * n + 2] = hg.kG[n];
    }
    for (int n = 0; n < nD; ++n) {
        int iG = hg.iD[n], jG = hg.jD[n], kG = hg.kD[n];
        if (!inb(iG, jG, kG) || !inb(hg.iDC[n], hg.jDC[n], hg.kDC[n]) ||
            !inb(hg.iDC[n] + 1, hg.jDC[n] + 1, hg.kDC[n] + 1))
            return fail(-1, "dead-cell or host-cell index out of range at n=" + std::to_string(n + 1));
        dT[n] = lin(iG, jG, kG);
        dB[n] = lin(hg.iDC[n], hg.jDC[n], hg.kDC[n]);
        dA[n] = (jG >= 1 && jG <= h->ny - 1 && kG >= 1 && kG <= h->nz - 1) ? 1 : 0;
        dS[n] = -1;
        double xIP = hg.ipD[3 * (size_t)n], yIP = hg.ipD[3 * (size_t)n + 1], zIP = hg.ipD[3 * (size_t)n + 2];
        double dist_int = 0.0;
        int sumk = 0;
        bool pure = false;
        const HostMet &m = h->hm[0];
        for (int r = 0; r < 8; ++r) {
            int ii = hg.iDC[n] + incI[r], jj = hg.jDC[n] + incJ[r], kk = hg.kDC[n] + incK[r];
            long q = F(ii, jj, kk);
            if (ii == iG && jj == jG && kk == kG) { dS[n] = r; }
            else {
                double dx = m.xc[ii] - xIP, dy = m.yc[jj] - yIP, dz = m.zc[kk] - zIP;
                double dist = dx * dx + dy * dy + dz * dz;
                if (h->h_iblank[q] == 0 || h->h_dead[q] == -1) {
                    sumk++;
                    dist_int = dist_int + 1.0 / dist;
                    dW[8 * (size_t)n + r] = 1.0 / dist;
                }
            }
            if (h->h_iblank[q] == 1 && h->h_dead[q] == 0) pure = true;
        }
        dI[n] = (pure && !(dist_int < 0.000001 || sumk == 1)) ? 1 : 0;
    }
    // dependency levels of the exact GCM pass: targets are the active dead
    // cells (list index n) and the ghost cells (nD + n); an edge runs from a
    // stencil cell that is itself a target to the target reading it
    std::vector<int> order;
    std::vector<int> level_off;
    bool sched_ok = false;
    {
        const long N1 = (long)(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
        std::vector<int> tid(N1, -1);
        int nt = nD + nG;
        for (int n = 0; n < nD; ++n) if (dA[n]) tid[F(hg.iD[n], hg.jD[n], hg.kD[n])] = n;
        for (int n = 0; n < nG; ++n) tid[F(hg.iG[n], hg.jG[n], hg.kG[n])] = nD + n;
        std::vector<std::vector<int>> users(nt);
        std::vector<int> indeg(nt, 0);
        auto add_deps = [&](int t, int bi, int bj, int bk, int si, int sj, int sk, bool none) {
            if (none) return;
            for (int r = 0; r < 8; ++r) {
                int ii = bi + incI[r], jj = bj + incJ[r], kk = bk + incK[r];
                if (ii == si && jj == sj && kk == sk) continue;
                int u = tid[F(ii, jj, kk)];
                if (u >= 0 && u != t) { users[u].push_back(t); indeg[t]++; }
            }
        };
        for (int n = 0; n < nD; ++n)
            if (dA[n]) add_deps(n, hg.iDC[n], hg.jDC[n], hg.kDC[n], hg.iD[n], hg.jD[n], hg.kD[n], dI[n] != 0);
        for (int n = 0; n < nG; ++n)
            add_deps(nD + n, hg.iC[n], hg.jC[n], hg.kC[n], hg.iG[n], hg.jG[n], hg.kG[n], false);
        std::vector<int> frontier, next;
        int active = 0;
        for (int t = 0; t < nt; ++t) {
            bool act = (t >= nD) || dA[t];
            if (!act) continue;
            active++;
            if (indeg[t] == 0) frontier.push_back(t);
        }
        level_off.push_back(0);
        while (!frontier.empty()) {
            std::sort(frontier.begin(), frontier.end());
            order.insert(order.end(), frontier.begin(), frontier.end());
            level_off.push_back((int)order.size());
            next.clear();
            for (int v : frontier)
                for (int u : users[v])
                    if (--indeg[u] == 0) next.push_back(u);
            frontier.swap(next);
        }
        sched_ok = ((int)order.size() == active);
    }

    // one buffer for everything
    size_t bytes = 0;
    auto add = [&](size_t b) { size_t o = bytes; bytes += (b + 255) / 256 * 256; return o; };
    size_t o_gT = add(8 * (size_t)nG), o_gB = add(8 * (size_t)nG), o_gS = add(4 * (size_t)nG),
           o_gC = add(64 * (size_t)nG), o_gA = add(8 * (size_t)nG), o_gP = add(8 * (size_t)nG),
           o_dT = add(8 * (size_t)nD), o_dB = add(8 * (size_t)nD), o_dS = add(4 * (size_t)nD),
           o_dA = add(4 * (size_t)nD), o_dI = add(4 * (size_t)nD), o_dC = add(64 * (size_t)nD),
           o_dW = add(64 * (size_t)nD), o_dL = add(8 * (size_t)nD), o_dP = add(8 * (size_t)nD),
           o_ijk = add(12 * (size_t)std::max(nG, 1)), o_off = add(64), o_vD = add(8 * (size_t)nD),
           o_vG = add(8 * (size_t)nG), o_rl = add(8 * (size_t)std::max(nG, 1)),
           o_node = add(sizeof(GcmNode) * std::max<size_t>(order.size(), 1)),
           o_gd = add(8 * std::max<size_t>(order.size(), 1));
    if (h->gbuf) { dfree(h, h->gbuf); h->gbuf = nullptr; }
    {
        unsigned char *gb = nullptr;
        int arc = dalloc(h, &gb, std::max<size_t>(bytes, 256));
        if (arc) return arc;
        h->gbuf = gb;
    }
    std::vector<unsigned char> hb(std::max<size_t>(bytes, 256), 0);
    auto cp = [&](size_t o, const void *src, size_t n) { if (n) memcpy(hb.data() + o, src, n); };
    std::vector<double> dpG(nG), dpD(nD), plD(nD);
    for (int n = 0; n < nG; ++n) dpG[n] = hg.dpdnG[n];
    for (int n = 0; n < nD; ++n) { dpD[n] = hg.dpdnD[n]; plD[n] = hg.plD[n]; }
    long off8[8];
    for (int r = 0; r < 8; ++r) off8[r] = incI[r] + d.sy * incJ[r] + d.sz * incK[r];
    cp(o_gT, gT.data(), 8 * (size_t)nG); cp(o_gB, gB.data(), 8 * (size_t)nG); cp(o_gS, gS.data(), 4 * (size_t)nG);
    cp(o_gC, hg.cN.data(), 64 * (size_t)nG); cp(o_gA, hg.alpha.data(), 8 * (size_t)nG); cp(o_gP, dpG.data(), 8 * (size_t)nG);
    cp(o_dT, dT.data(), 8 * (size_t)nD); cp(o_dB, dB.data(), 8 * (size_t)nD); cp(o_dS, dS.data(), 4 * (size_t)nD);
    cp(o_dA, dA.data(), 4 * (size_t)nD); cp(o_dI, dI.data(), 4 * (size_t)nD); cp(o_dC, hg.cDN.data(), 64 * (size_t)nD);
    cp(o_dW, dW.data(), 64 * (size_t)nD); cp(o_dL, plD.data(), 8 * (size_t)nD); cp(o_dP, dpD.data(), 8 * (size_t)nD);
    cp(o_ijk, gijk.data(), 12 * (size_t)nG); cp(o_off, off8, 64);
    // the schedule's relations as records in schedule order (dead cells are
    // list entries < nD, ghost cells nD + n)
    {
        std::vector<GcmNode> nodes(order.size());
        for (size_t t = 0; t < order.size(); ++t) {
            GcmNode &nd = nodes[t];
            const int e = order[t];
            if (e < nD) {
                const int n = e;
                nd.target = (int)dT[n];
                nd.base = (int)dB[n];
                nd.self = dS[n];
                nd.kind = dI[n] ? 2 : 1;
                for (int r = 0; r < 8; ++r) nd.coef[r] = dI[n] ? dW[8 * (size_t)n + r] : hg.cDN[8 * (size_t)n + r];
                nd.dpdn = dpD[n];
                nd.pl = plD[n];
            } else {
                const int n = e - nD;
                nd.target = (int)gT[n];
                nd.base = (int)gB[n];
                nd.self = gS[n];
                nd.kind = 0;
                for (int r = 0; r < 8; ++r) nd.coef[r] = hg.cN[8 * (size_t)n + r];
                nd.dpdn = dpG[n];
                nd.pl = 0.0;
            }
        }
        cp(o_node, nodes.data(), sizeof(GcmNode) * nodes.size());
    }
    HIPCHK(hipMemcpyAsync(h->gbuf, hb.data(), bytes, hipMemcpyHostToDevice, h->st));
    unsigned char *B = (unsigned char *)h->gbuf;
    GcmDev &g = h->g;
    g.nG = nG; g.nD = nD;
    g.gTarget = (const long *)(B + o_gT); g.gBase = (const long *)(B + o_gB); g.gSelf = (const int *)(B + o_gS);
    g.gCoef = (const double *)(B + o_gC); g.gAlpha = (const double *)(B + o_gA); g.gDpdn = (const double *)(B + o_gP);
    g.dTarget = (const long *)(B + o_dT); g.dBase = (const long *)(B + o_dB); g.dSelf = (const int *)(B + o_dS);
    g.dActive = (const int *)(B + o_dA); g.dIdw = (const int *)(B + o_dI); g.dCoef = (const double *)(B + o_dC);
    g.dW = (const double *)(B + o_dW); g.dPl = (const double *)(B + o_dL); g.dDpdn = (const double *)(B + o_dP);
    h->gijk = (int *)(B + o_ijk);
    h->off8 = (long *)(B + o_off);
    h->gvalD = (double *)(B + o_vD);
    h->gvalG = (double *)(B + o_vG);
    h->reslap = (double *)(B + o_rl);
    h->gnode = (GcmNode *)(B + o_node);
    h->gdelta = (double *)(B + o_gd);
    h->glevel_off = level_off;
    h->gsched_ok = sched_ok;
    h->gsched_levels = (int)level_off.size() - 1;
    long need = std::max<long>((nG + 255) / 256, (nD + 255) / 256);
    if (need > h->npartial) return fail(-1, "GCM list larger than the reduction scratch");
    HIPCHK(hipStreamSynchronize(h->st));
    h->gcm_dirty = false;
    return 0;
}

// ---------------------------------------------------------------------------
void timing_flush(fc_pson *h)
{
    if (h->ev_used == 0) return;
    hipEventSynchronize(h->ev[h->ev_used - 1]);
    for (size_t i = 0; i + 1 < h->ev_used; i += 2) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]);
        h->sweep_ms += ms;
        h->sweep_launches++;
    }
    h->ev_used = 0;
}

hipEvent_t next_event(fc_pson *h)
{
    if (h->ev_used == h->ev.size()) {
        if (h->ev.size() >= 4096) timing_flush(h);
        if (h->ev_used == h->ev.size()) {
            hipEvent_t e;
            hipEventCreate(&e);
            h->ev.push_back(e);
        }
    }
    return h->ev[h->ev_used++];
}

// The masks one V-cycle runs with: the reference's (igk held rows, the ibk
// volume branches of the transfers) or, for the correction cycles with
// bodies as walls, the wall masks (no volume branches; renormalised
// prolongation next to held coarse cells).
struct VcMasks {
    const std::vector<int8_t *> *row, *ibk, *pflag;
    const std::vector<int8_t *> *rflag;   // null: no restriction volume branch
    int walls;
};
VcMasks vc_masks(const fc_pson *h, bool walls)
{
    if (walls) return VcMasks{&h->wmask, &h->wmask, &h->wpflag, nullptr, 1};
    return VcMasks{&h->igk, &h->ibk, &h->pflag, &h->rflag, 0};
}

// Coarsest level of the wall correction cycles solved directly in FP64
// (coarse_solve = 1, at most FC_COARSE_DIRECT_MAX free cells).  The
// unknowns are the free cells; a held cell keeps e = 0.  Row t is the wall
// operator sum over open faces c_f (e_f - e_t) (a wall face carries nothing;
// an outer face follows the homogeneous boundary fill of a coarse level:
// Neumann ghost = e_t adds nothing, Dirichlet ghost = -e_t adds -2 c_f e_t).
// A group of connected free cells that touches no Dirichlet face is
// singular: its solvable right-hand sides have zero volume-weighted sum
// (the volume-scaled operator is symmetric), so the right-hand side is
// projected onto them (its volume-weighted mean over the group removed)
// and the group's first cell is pinned to e = 0.  M = A_pinned^-1 Z P (P
// the projection, Z zeroing the pinned rows) is formed once on the host by
// LU with partial pivoting and applied on the device as a dense product.
int ensure_coarse_direct(fc_pson *h)
{
    if (h->cd_state != 0) return 0;
    h->cd_state = -1;
    if (h->prm.coarse_solve != 1 || h->prm.vc_walls < 1 || !h->have_masks) return 0;
    if (bcinfo(h, 0, true).per) return 0;     // periodic faces: the line-SOR pass stays
    const int l = h->nlev - 1;
    const LevelDims &c = h->dims[l];
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const size_t nf = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    std::vector<int8_t> wm(nf);
    int rc = to_host(h, wm.data(), h->wmask[l], l);
    if (rc) return rc;
    auto F = [&](int i, int j, int k) { return (size_t)i + (size_t)(nx + 2) * ((size_t)j + (size_t)(ny + 2) * k); };
    std::vector<int> id(nf, -1);
    std::vector<int> ci, cj, ck;
    std::vector<long> q;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j)
            for (int i = 1; i <= nx - 1; ++i)
                if (!(wm[F(i, j, k)] & 1)) {
                    id[F(i, j, k)] = (int)q.size();
                    ci.push_back(i); cj.push_back(j); ck.push_back(k);
                    q.push_back((long)i + (long)c.px * ((long)j + (long)(ny + 2) * k));
                }
    const int n = (int)q.size();
    if (n == 0 || n > FC_COARSE_DIRECT_MAX) return 0;
    const HostMet &m = h->hm[l];
    const size_t N = (size_t)n;
    std::vector<double> A(N * N, 0.0), V(N);
    std::vector<int> dirich(N, 0);
    for (int t = 0; t < n; ++t) {
        const int i = ci[t], j = cj[t], k = ck[t];
        const int8_t b = wm[F(i, j, k)];
        const double cf[6] = {m.dxcinv[i] * m.dxinv[i], m.dxcinv[i + 1] * m.dxinv[i], m.dycinv[j] * m.dyinv[j],
                              m.dycinv[j + 1] * m.dyinv[j], m.dzcinv[k] * m.dzinv[k], m.dzcinv[k + 1] * m.dzinv[k]};
        const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
        for (int f = 0; f < 6; ++f) {
            if (b & (2 << f)) continue;                          // wall face
            const int ii = i + di[f], jj = j + dj[f], kk = k + dk[f];
            const bool outer = ii == 0 || ii == nx || jj == 0 || jj == ny || kk == 0 || kk == nz;
            if (outer) {
                if (h->prm.pbc[f] == FC_PBC_DIRICHLET) { A[N * t + t] -= 2.0 * cf[f]; dirich[t] = 1; }
                continue;
            }
            const int sidx = id[F(ii, jj, kk)];
            if (sidx < 0) return fail(-1, "coarse direct solve: open face towards a held cell");
            A[N * t + sidx] += cf[f];
            A[N * t + t] -= cf[f];
        }
        V[t] = 1.0 / (m.dxinv[i] * m.dyinv[j] * m.dzinv[k]);
    }
    // groups of connected free cells
    std::vector<int> comp(N, -1), stack;
    std::vector<int> pins;                 // first cell of each floating group
    std::vector<std::vector<int>> members;
    for (int t0 = 0; t0 < n; ++t0) {
        if (comp[t0] >= 0) continue;
        const int g = (int)members.size();
        members.emplace_back();
        bool fixed = false;
        comp[t0] = g;
        stack.assign(1, t0);
        while (!stack.empty()) {
            const int t = stack.back();
            stack.pop_back();
            members[g].push_back(t);
            fixed = fixed || dirich[t];
            const int i = ci[t], j = cj[t], k = ck[t];
            const int8_t b = wm[F(i, j, k)];
            const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
            for (int f = 0; f < 6; ++f) {
                if (b & (2 << f)) continue;
                const int ii = i + di[f], jj = j + dj[f], kk = k + dk[f];
                if (ii == 0 || ii == nx || jj == 0 || jj == ny || kk == 0 || kk == nz) continue;
                const int u = id[F(ii, jj, kk)];
                if (u >= 0 && comp[u] < 0) { comp[u] = g; stack.push_back(u); }
            }
        }
        if (fixed) members[g].clear();    // regular: no projection, no pin
        else pins.push_back(*std::min_element(members[g].begin(), members[g].end()));
    }
    for (int t : pins) {
        for (size_t s2 = 0; s2 < N; ++s2) A[N * t + s2] = 0.0;
        A[N * t + t] = 1.0;
    }
    // LU with partial pivoting, then X = A^-1 column by column
    std::vector<int> piv(N);
    for (size_t col = 0; col < N; ++col) {
        size_t pr = col;
        double best = std::fabs(A[N * col + col]);
        for (size_t r = col + 1; r < N; ++r)
            if (std::fabs(A[N * r + col]) > best) { best = std::fabs(A[N * r + col]); pr = r; }
        if (!(best > 0.0)) return fail(-1, "coarse direct solve: singular coarsest operator");
        piv[col] = (int)pr;
        if (pr != col)
            for (size_t c2 = 0; c2 < N; ++c2) std::swap(A[N * col + c2], A[N * pr + c2]);
        const double inv = 1.0 / A[N * col + col];
        for (size_t r = col + 1; r < N; ++r) {
            const double f = A[N * r + col] * inv;
            A[N * r + col] = f;
            if (f != 0.0)
                for (size_t c2 = col + 1; c2 < N; ++c2) A[N * r + c2] -= f * A[N * col + c2];
        }
    }
    std::vector<double> X(N * N, 0.0), y(N);
    for (size_t e = 0; e < N; ++e) {
        for (size_t r = 0; r < N; ++r) y[r] = (r == e) ? 1.0 : 0.0;
        for (size_t r = 0; r < N; ++r) std::swap(y[r], y[piv[r]]);
        for (size_t r = 0; r < N; ++r) {
            double a = y[r];
            for (size_t c2 = 0; c2 < r; ++c2) a -= A[N * r + c2] * y[c2];
            y[r] = a;
        }
        for (size_t r = N; r-- > 0;) {
            double a = y[r];
            for (size_t c2 = r + 1; c2 < N; ++c2) a -= A[N * r + c2] * y[c2];
            y[r] = a / A[N * r + r];
        }
        for (size_t r = 0; r < N; ++r) X[N * r + e] = y[r];
    }
    // M = X Z P: pinned columns dropped, then the projection of each floating
    // group (volume-weighted mean removed)
    for (int t : pins)
        for (size_t r = 0; r < N; ++r) X[N * r + t] = 0.0;
    for (const auto &mem : members) {
        if (mem.empty()) continue;
        double vs = 0.0;
        for (int t : mem) vs += V[t];
        for (size_t r = 0; r < N; ++r) {
            double xu = 0.0;
            for (int t : mem) xu += X[N * r + t];
            if (xu == 0.0) continue;
            for (int t : mem) X[N * r + t] -= xu * V[t] / vs;
        }
    }
    if ((rc = dalloc(h, &h->cd_M, N * N)) || (rc = dalloc(h, &h->cd_q, N))) return rc;
    HIPCHK(hipMemcpyAsync(h->cd_M, X.data(), sizeof(double) * N * N, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->cd_q, q.data(), sizeof(long) * N, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->cd_n = n;
    h->cd_state = 1;
    return 0;
}

void drop_coarse_direct(fc_pson *h)
{
    if (h->st) hipStreamSynchronize(h->st);
    dfree(h, h->cd_M);
    dfree(h, h->cd_q);
    h->cd_M = nullptr;
    h->cd_q = nullptr;
    h->cd_n = 0;
    h->cd_state = 0;
}

template <typename T>
// fill = false: the outer ghosts already hold what the boundary fill would
// set (a coarse level right after the fused restriction into it)
void smooth(fc_pson *h, std::vector<LevelPtrs<T>> &L, const std::vector<DevMet<T>> &M, int l, bool homo1,
            bool fill = true, bool walls = false)
{
    const VcMasks vm = vc_masks(h, walls);
    const LevelDims &d = h->dims[l];
    BcInfo bc = bcinfo(h, l, homo1);
    // set_outer_pressure_bc before the first half sweep; each half sweep then
    // keeps the ghosts of the cells it updates current (the reference fills
    // them before every direction, CORE_MG.F90:640-645)
    if (fill) bc_fill<T>(h->st, d, L[l].phi, bc);
    for (int dir = 0; dir < 3; ++dir)
        for (int c = 0; c < 2; ++c) {
            const bool t = h->timing && l == 0;
            if (t) hipEventRecord(next_event(h), h->st);
            sweep<T>(h->st, d, M[l].m, L[l].phi, L[l].rhs, (*vm.row)[l], dir, c, h->prm.omega_pson, bc);
            if (t) hipEventRecord(next_event(h), h->st);
            // periodic ghosts depend on both ends of a line of cells: refresh
            if (bc.per) bc_fill<T>(h->st, d, L[l].phi, bc);
        }
}

// V-cycle shape in use (the cycle parameter, with its automatic choice)
int cycle_shape(const fc_pson *h)
{
    return (h->prm.cycle == 2) ? 2 : 1;
}

// one V-cycle of mg_solve (CORE_MG.F90:1618-1766) in the GPU ordering
template <typename T>
// fresh: level 1's phi is zero with its outer ghosts (the FP32 correction
// right after resid64_to32), so its first boundary fill is skipped
// walls: the correction cycle with bodies as walls (vc_masks)
void vcycle(fc_pson *h, std::vector<LevelPtrs<T>> &L, const std::vector<DevMet<T>> &M, bool homo1,
            bool fresh = false, bool walls = false)
{
    const int ml = h->nlev;
    const VcMasks vm = vc_masks(h, walls);
    for (int l = 0; l < ml; ++l) {
        const LevelDims &d = h->dims[l];
        // (level l > 0 had its phi and outer ghosts zeroed by the restriction
        // into it)
        if (walls && l == ml - 1 && ml > 1 && h->cd_state == 1) {
            coarse_direct<T>(h->st, h->cd_n, h->cd_M, h->cd_q, L[l].rhs, L[l].phi);
            bc_fill<T>(h->st, d, L[l].phi, bcinfo(h, l, homo1));
        } else {
            smooth<T>(h, L, M, l, homo1, l == 0 && !fresh, walls);
        }
        if (l != ml - 1)
            resrestrict<T>(h->st, d, h->dims[l + 1], M[l].m, M[l + 1].m, L[l].phi, L[l].rhs, (*vm.row)[l],
                           (*vm.ibk)[l], vm.rflag ? (*vm.rflag)[l + 1] : nullptr, L[l + 1].rhs, L[l + 1].phi);
    }
    for (int l = ml - 1; l >= 1; --l) {
        if (l != ml - 1) smooth<T>(h, L, M, l, homo1, true, walls);
        prolong<T>(h->st, h->dims[l - 1], h->dims[l], M[l - 1].m, M[l].m, L[l].phi, L[l - 1].phi, (*vm.ibk)[l - 1],
                   (*vm.ibk)[l], (*vm.pflag)[l], vm.walls);
    }
    // cycle shape 2: one post-smoothing pass on the finest level (the
    // reference smooths the finest level only on the way down)
    if (cycle_shape(h) == 2) smooth<T>(h, L, M, 0, homo1, true, walls);
}

int read_scalars(fc_pson *h, int n)
{
    HIPCHK(hipMemcpyAsync(h->hscal, h->scal, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int read_scalars_at(fc_pson *h, int off, int n)
{
    HIPCHK(hipMemcpyAsync(h->hscal + off, h->scal + off, sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

double node_norm(const fc_pson *h) { return (double)((long)h->nx * h->ny * h->nz); }

// GCM_ghostcell_pressure on the resident pressure
int gcm_update(fc_pson *h, fc_pson_stats *st, double *gdelta = nullptr)
{
    int rc = finalize_gcm(h);
    if (rc) return rc;
    int iterGC = 0;
    double resPresMax = 1.0e10;
    const int homo = h->prm.is_PBC_homogeneous == 1;
    if (h->gsched_ok && h->prm.mix_GC_form != 1 && h->prm.itermax_gc > 0) {
        // one exact pass in dependency order = the converged list-order sweeps
        gcm_levels(h->st, h->gnode, h->glevel_off.data(), h->gsched_levels, h->p, h->pprev, h->off8, homo,
                   gdelta);
        if (st) { st->gcm_sweeps++; st->gcm_change = -1.0; }
        return 0;
    }
    while (iterGC < h->prm.itermax_gc && resPresMax > h->prm.restol_gcp) {
        gcm_dead(h->st, h->g, h->p, h->pprev, h->off8, homo, h->gvalD);
        gcm_scatter(h->st, h->g.nD, h->g.dTarget, h->g.dActive, h->gvalD, h->p, h->partial, h->scal + 10);
        gcm_ghost(h->st, h->g, h->p, h->div, h->off8, homo, h->prm.mix_GC_form == 1, h->dims[0], h->l1met, h->gijk,
                  h->gvalG, h->reslap);
        gcm_scatter(h->st, h->g.nG, h->g.gTarget, nullptr, h->gvalG, h->p, h->partial, h->scal + 11);
        HIPCHK(hipMemcpyAsync(h->hscal + 10, h->scal + 10, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        resPresMax = std::max(h->hscal[10], h->hscal[11]);
        iterGC++;
        if (st) st->gcm_sweeps++;
    }
    if (st) st->gcm_change = resPresMax;
    return 0;
}

// itsolv (CORE_POISSON.F90:205-456) on the resident pressure and rhs: one
// plain line-SOR pass, x-, y- then z-lines, each direction as the two
// checkerboard colours (the GPU ordering; the reference's is lexicographic
// for x and z and zebra for y), relaxation var + omega (x - var) on the
// interior rows, masked rows (flag = 0) held, the ghost layers read as they
// are (the caller's outer-BC fill)
void itsolv_resident(fc_pson *h)
{
    const LevelDims &d0 = h->dims[0];
    BcInfo bc = bcinfo(h, 0, false);
    for (int dir = 0; dir < 3; ++dir)
        for (int c = 0; c < 2; ++c)
            sweep_itsolv(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], dir, c, h->prm.omega_pson, bc);
}

int mg_solve_resident(fc_pson *h, fc_pson_stats *st, double *resm_out);

// the inner solve of the outer loop: mg_solve or, for it_solver_type = 1,
// itsolv (CORE_POISSON.F90:170-176)
int inner_solve(fc_pson *h, fc_pson_stats *st, double *resm_out)
{
    if (h->prm.it_solver_type == FC_IT_SOLVER_TYPE_LSOR) {
        itsolv_resident(h);
        return 0;
    }
    return mg_solve_resident(h, st, resm_out);
}

// mg_solve on the resident pressure
int mg_solve_resident(fc_pson *h, fc_pson_stats *st, double *resm_out)
{
    const LevelDims &d0 = h->dims[0];
    int iter_mg = 0;
    double resm = 1.0e6;
    if (h->prm.precision == FC_PREC_MIXED) {
        int rc = ensure_mixed(h);
        if (rc) return rc;
        std::vector<DevMet<float>> &M32 = h->m32;
        BcInfo bc = bcinfo(h, 0, false);
        bc_fill<double>(h->st, d0, h->p, bc);
        const bool need_res = h->prm.restol_pson > 0.0 || resm_out;
        resid64_to32(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->L32[0].rhs, h->L32[0].phi, h->partial,
                     need_res ? h->scal : nullptr);
        if (need_res) {
            if ((rc = read_scalars(h, 2))) return rc;
            resm = std::sqrt(h->hscal[0] / node_norm(h));
        }
        while (iter_mg <= h->prm.iterInter && resm > h->prm.restol_pson) {
            int nc = std::max(1, h->prm.fp32_cycles);
            for (int c = 0; c < nc; ++c) {
                vcycle<float>(h, h->L32, M32, true, c == 0);
                if (st) st->vcycles_fp32++;
            }
            add_correction(h->st, d0.n, h->p, h->L32[0].phi);
            bc_fill<double>(h->st, d0, h->p, bc);
            if (st) st->refine_steps++;
            iter_mg++;
            // the residual only steers the loop: skip it after the last step
            if (iter_mg > h->prm.iterInter && !resm_out) break;
            resid64_to32(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->L32[0].rhs, h->L32[0].phi, h->partial,
                         need_res ? h->scal : nullptr);
            if (need_res) {
                if ((rc = read_scalars(h, 2))) return rc;
                resm = std::sqrt(h->hscal[0] / node_norm(h));
            }
        }
    } else {
        while (iter_mg <= h->prm.iterInter && resm > h->prm.restol_pson) {
            vcycle<double>(h, h->L64, h->m64, false);
            resnorm<double>(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->partial, h->scal);
            int rc = read_scalars(h, 2);
            if (rc) return rc;
            resm = std::sqrt(h->hscal[0] / node_norm(h));
            if (h->prm.infoconv == 1) printf("MG: residual check at lev 1: %4d  %12.5E\n", iter_mg, resm);
            iter_mg++;
            if (st) st->vcycles_fp64++;
        }
    }
    if (resm_out) *resm_out = resm;
    return 0;
}

int prepare_gvel(fc_pson *h);
size_t velbc_len(const fc_pson *h, int a);
double *vbc_ptr(fc_pson *h, int a, int c);

// bit f set when outer face f (x1, x2, y1, y2, z1, z2) is zero-gradient
int zg_mask(const fc_pson *h)
{
    int m = 0;
    for (int f = 0; f < 6; ++f)
        if (h->prm.bc_vel[f] == FC_BC_TYPE_ZERO_GRADIENT) m |= 1 << f;
    return m;
}

// GCM_enforce_p_compatibility (GCM_GHOSTCELL_UPDATE.F90:865-1003) on the
// resident pressure: with a zero-gradient face and no Dirichlet pressure
// face, the net pressure-gradient flux of the flag cells (and of the ghost
// cells less res_laplacian when mix_GC_form = 1) is removed through the
// pgrad planes of the zero-gradient faces.  All on the device.
int enforce_p_compat(fc_pson *h)
{
    const int zg = zg_mask(h);
    if (!zg) return 0;
    for (int f = 0; f < 6; ++f)
        if (h->prm.pbc[f] == FC_PBC_DIRICHLET) return 0;
    const bool mix = h->prm.mix_GC_form == 1;
    if (mix && h->gvel_dirty) {
        int rc = prepare_gvel(h);
        if (rc) return rc;
    }
    pressure_flux(h->st, h->dims[0], h->p, h->igk[0], h->l1met, h->l1rhs, h->gvel, mix ? h->ngvel : 0,
                  mix ? h->reslap : nullptr, h->partial, h->scal + 114);
    pgrad_correct(h->st, h->dims[0], zg, h->scal + 114, h->outflow_area, h->pg);
    return 0;
}

// One pass of the solve_poisson loop body (CORE_POISSON.F90:129-191): mean
// pinning, outer BC, p_temp, GCM update, mg_solve, calc_residual and
// calc_residual2.  Returns maxres, maxres2 and the flag-cell relative
// residual through *out (3 values).
// The accelerated loop pins its iterate to p(1,1,1) = 0 when every pressure
// face is Neumann (the solution is defined up to a constant; the reference
// loop pins its mean then, CORE_POISSON.F90:129-150); with a Dirichlet face
// the solution is unique and nothing is pinned: -1.
long pin_index(const fc_pson *h)
{
    for (int f = 0; f < 6; ++f)
        if (h->prm.pbc[f] != FC_PBC_NEUMANN) return -1;
    return (long)1 + h->dims[0].sy + h->dims[0].sz;
}

int outer_step(fc_pson *h, fc_pson_stats &st, double *out)
{
    const LevelDims &d0 = h->dims[0];
    bool allN = true;
    for (int f = 0; f < 6; ++f) if (h->prm.pbc[f] != FC_PBC_NEUMANN) allN = false;
    BcInfo bc = bcinfo(h, 0, false);
    int rc;
    if (allN) pin(h->st, d0, h->p, h->scal + 20);
    bc_fill<double>(h->st, d0, h->p, bc);
    if ((rc = enforce_p_compat(h))) return rc;
    HIPCHK(hipMemcpyAsync(h->ptemp, h->p, sizeof(double) * d0.n, hipMemcpyDeviceToDevice, h->st));
    if ((rc = gcm_update(h, &st))) return rc;
    if ((rc = inner_solve(h, &st, nullptr))) return rc;
    resnorm<double>(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->partial, h->scal);
    maxdiff(h->st, d0, h->p, h->ptemp, h->ibk[0], h->partial + h->npartial, h->scal + 2);
    if ((rc = read_scalars(h, 3))) return rc;
    out[0] = std::sqrt(h->hscal[0] / node_norm(h));
    out[1] = h->hscal[2];
    out[2] = (h->hscal[1] > 0.0) ? std::sqrt(h->hscal[0] / h->hscal[1]) : std::sqrt(h->hscal[0]);
    return 0;
}

int ensure_aa(fc_pson *h, int m)
{
    if (h->aa_m == m) return 0;
    if (h->aa_m != 0) return fail(-1, "Anderson depth cannot change on a handle");
    const long n = h->dims[0].n;
    int rc;
    if ((rc = dalloc(h, &h->aa_fprev, n)) || (rc = dalloc(h, &h->aa_c, n)) || (rc = dalloc(h, &h->aa_ring, 16)))
        return rc;
    for (int i = 0; i < 8; ++i) { h->aa_cols.dF[i] = nullptr; h->aa_cols.dG[i] = nullptr; }
    for (int i = 0; i < m; ++i)
        if ((rc = dalloc(h, &h->aa_cols.dF[i], n)) || (rc = dalloc(h, &h->aa_cols.dG[i], n))) return rc;
    int ring[16] = {m, 0, 0, 0};
    HIPCHK(hipMemcpyAsync(h->aa_ring, ring, sizeof(ring), hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->aa_m = m;
    return 0;
}

// Defect correction with the map's step kept in e (no G or x field): one
// refinement step per outer iteration and the exact GCM pass (or none), so
// that G - x = e + (GCM changes) - x(1,1,1).  The mixed path always runs
// this way when it can (FP32 e); the FP64 path does with the wall correction
// cycles (FP64 e), and otherwise runs its V-cycles on p itself as the
// reference does.
bool mixed_fold(const fc_pson *h)
{
    const fc_pson_params &P = h->prm;
    return P.it_solver_type == FC_IT_SOLVER_TYPE_MG && P.iterInter == 0 && P.restol_pson <= 0.0 &&
           (P.itermax_gc <= 0 || (h->gsched_ok && P.mix_GC_form != 1)) &&
           (P.precision == FC_PREC_MIXED || P.vc_walls >= 1);
}

// the FP64 correction hierarchy of the FP64 fold: level 1 holds e and its
// right-hand side in fields of its own, the coarse levels are the FP64
// levels
int ensure_corr64(fc_pson *h)
{
    if (!h->L64c.empty()) return 0;
    int rc;
    double *e = nullptr, *r = nullptr;
    if ((rc = dalloc(h, &e, h->dims[0].n)) || (rc = dalloc(h, &r, h->dims[0].n))) {
        dfree(h, e);
        return rc;
    }
    HIPCHK(hipStreamSynchronize(h->st));
    h->L64c = h->L64;
    h->L64c[0].phi = e;
    h->L64c[0].rhs = r;
    return 0;
}

// one outer iteration of the fold: FP64 residual of p as the right-hand
// side of the correction, V-cycles on e (TE = float: the mixed path's FP32
// hierarchy; double: the FP64 correction hierarchy), the GCM changes moved
// into e, the Anderson history and mixing coefficients
template <typename TE>
int fold_body(fc_pson *h, std::vector<LevelPtrs<TE>> &L, const std::vector<DevMet<TE>> &M, int *vc)
{
    const LevelDims &d0 = h->dims[0];
    const long q111 = pin_index(h);
    TE *e = L[0].phi;
    resid64_to32_sums<TE>(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], L[0].rhs, e, h->partial);
    const bool walls = h->prm.vc_walls >= 1;
    for (int c = 0; c < std::max(1, h->prm.fp32_cycles); ++c) {
        vcycle<TE>(h, L, M, true, c == 0, walls);
        (*vc)++;
    }
    if (h->prm.itermax_gc > 0) gcm_fold<TE>(h->st, h->gnode, h->glevel_off.back(), h->gdelta, e, h->p, walls);
    aa_fresid<TE>(h->st, d0, e, h->aa_fprev, h->aa_c, h->aa_cols, h->aa_ring, h->partial, h->scal);
    aa_lsq<TE>(h->st, h->aa_ring, h->scal, h->aa_cols, h->p, e, q111, 0);
    return 0;
}

// the iterate x (FP64 path, or mixed without the fold) and G (mixed without
// the fold) on first use
int ensure_aa_fields(fc_pson *h)
{
    const long n = h->dims[0].n;
    const bool fold = mixed_fold(h), mixed = h->prm.precision == FC_PREC_MIXED;
    int rc;
    if (!fold && !h->aa_x && (rc = dalloc(h, &h->aa_x, n))) return rc;
    if (mixed && !fold && !h->aa_g && (rc = dalloc(h, &h->aa_g, n))) return rc;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

// One evaluation of the solve_poisson loop body G(x) for the accelerated
// loop (CORE_POISSON.F90:129-191): the caller has written the pinned iterate
// into p.  Outer BC, GCM update, mg_solve; then the residual sums of G, the
// Anderson history update and the mixing coefficients, all on the device.
// G is p (FP64) or p + e (mixed).  No host synchronisation unless
// restol_pson > 0 or the GCM update needs its iterative form.
int aa_body(fc_pson *h, const double *G, int *vc64, int *vc32, int *gsw)
{
    const LevelDims &d0 = h->dims[0];
    BcInfo bc = bcinfo(h, 0, false);
    int rc;
    bc_fill<double>(h->st, d0, h->p, bc);
    if ((rc = enforce_p_compat(h))) return rc;
    fc_pson_stats gst;
    memset(&gst, 0, sizeof(gst));
    const bool fold = mixed_fold(h);
    if ((rc = gcm_update(h, &gst, fold ? h->gdelta : nullptr))) return rc;
    *gsw += gst.gcm_sweeps;
    const long q111 = pin_index(h);
    if (fold) {
        // one refinement step; the map's step G - p stays in e
        if (h->prm.precision == FC_PREC_MIXED) return fold_body<float>(h, h->L32, h->m32, vc32);
        return fold_body<double>(h, h->L64c, h->m64, vc64);
    }
    const int ncyc = h->prm.iterInter + 1;
    if (h->prm.precision == FC_PREC_MIXED) {
        const int nc = std::max(1, h->prm.fp32_cycles);
        for (int it = 0; it < ncyc; ++it) {
            if (it > 0) {
                add_correction(h->st, d0.n, h->p, h->L32[0].phi);
                bc_fill<double>(h->st, d0, h->p, bc);
            }
            resid64_to32(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->L32[0].rhs, h->L32[0].phi, h->partial,
                         h->prm.restol_pson > 0.0 ? h->scal + 100 : nullptr);
            if (h->prm.restol_pson > 0.0) {
                if ((rc = read_scalars_at(h, 100, 1))) return rc;
                if (std::sqrt(h->hscal[100] / node_norm(h)) <= h->prm.restol_pson) break;
            }
            for (int c = 0; c < nc; ++c) {
                vcycle<float>(h, h->L32, h->m32, true, c == 0);
                (*vc32)++;
            }
        }
        add_into(h->st, d0, h->p, h->L32[0].phi, h->aa_g);
        bc_fill<double>(h->st, d0, h->aa_g, bc);
    } else if (h->prm.it_solver_type == FC_IT_SOLVER_TYPE_LSOR) {
        itsolv_resident(h);
    } else {
        for (int it = 0; it < ncyc; ++it) {
            if (h->prm.restol_pson > 0.0) {
                resnorm<double>(h->st, d0, h->m64[0].m, h->p, h->div, h->igk[0], h->partial, h->scal + 100);
                if ((rc = read_scalars_at(h, 100, 1))) return rc;
                if (std::sqrt(h->hscal[100] / node_norm(h)) <= h->prm.restol_pson) break;
            }
            vcycle<double>(h, h->L64, h->m64, false);
            (*vc64)++;
        }
    }
    aa_resid(h->st, d0, h->m64[0].m, G, h->div, h->igk[0], h->aa_x, h->aa_fprev, h->aa_c, h->aa_cols, h->aa_ring,
             h->partial, h->scal);
    aa_lsq<float>(h->st, h->aa_ring, h->scal, h->aa_cols, G, nullptr, q111, 0);
    return 0;
}

// solve_poisson (CORE_POISSON.F90:83-200) on the resident buffers
int solve_resident(fc_pson *h, fc_pson_stats *st_out)
{
    fc_pson_stats st;
    memset(&st, 0, sizeof(st));
    if (h->prm.bc_vel[4] == FC_BC_TYPE_PERIODIC && h->prm.bc_vel[5] == FC_BC_TYPE_PERIODIC)
        return fail(-2, "periodic condition in z is not implemented (as in the reference, "
                        "CORE_SET_OUTER_BC.F90:1028-1038)");
    if (h->prm.it_solver_type != FC_IT_SOLVER_TYPE_MG && h->prm.it_solver_type != FC_IT_SOLVER_TYPE_LSOR)
        return fail(-1, "it_solver_type must be 1 (LSOR) or 3 (MG)");
    if (h->prm.it_solver_type == FC_IT_SOLVER_TYPE_LSOR && h->prm.precision != FC_PREC_FP64)
        return fail(-2, "itsolv (it_solver_type = 1) runs in FP64 only");
    int rc = finalize_gcm(h);
    if (rc) return rc;
    if (h->prm.precision == FC_PREC_MIXED && (rc = ensure_mixed(h))) return rc;
    // the ghost list of the compatibility flux and the coarsest-level
    // operator are uploaded before any capture
    if (zg_mask(h) && h->prm.mix_GC_form == 1 && h->gvel_dirty && (rc = prepare_gvel(h))) return rc;
    if ((rc = ensure_coarse_direct(h))) return rc;
    // events and the iteration graph are released on every exit path
    struct Owned {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        ~Owned()
        {
            if (gexec) hipGraphExecDestroy(gexec);
            if (graph) hipGraphDestroy(graph);
            if (e0) hipEventDestroy(e0);
            if (e1) hipEventDestroy(e1);
        }
    } own;
    HIPCHK(hipEventCreate(&own.e0));
    HIPCHK(hipEventCreate(&own.e1));
    hipEvent_t e0 = own.e0, e1 = own.e1;
    hipEventRecord(e0, h->st);
    const LevelDims &d0 = h->dims[0];
    int iter = 0;
    double r3[3] = {1.0e10, 1.0e10, 1.0};
    int converged = 0;
    if (h->prm.outer_rule != FC_OUTER_RELRES) {
        // the reference's loop: stop on max|p - p_temp| <= restol_gcp
        while (iter < h->prm.itermax_pson && r3[1] > h->prm.restol_gcp) {
            if ((rc = outer_step(h, st, r3))) return rc;
            iter++;
        }
        converged = (r3[1] <= h->prm.restol_gcp);
    } else {
        // Anderson-accelerated fixed-point iteration of the same loop body
        int m = h->prm.aa_depth > 0 ? std::min(h->prm.aa_depth, 8) : 8;
        if ((rc = ensure_aa(h, m))) return rc;
        const bool mixed = h->prm.precision == FC_PREC_MIXED;
        if (mixed && (rc = ensure_mixed(h))) return rc;
        if ((rc = finalize_gcm(h)) || (rc = ensure_aa_fields(h))) return rc;
        const bool fold = mixed_fold(h);
        if (fold && !mixed && (rc = ensure_corr64(h))) return rc;
        const long q111 = pin_index(h);
        // the map value: G (FP64: p itself; mixed: aa_g), or p + e with the fold
        const double *G = (mixed && !fold) ? h->aa_g : h->p;
        const float *E = (fold && mixed) ? h->L32[0].phi : nullptr;
        const double *E64 = (fold && !mixed) ? h->L64c[0].phi : nullptr;
        // x_0 = the incoming pressure: reset the ring, pin it
        aa_lsq<float>(h->st, h->aa_ring, h->scal, h->aa_cols, h->p, nullptr, q111, 1);
        aa_next<float>(h->st, d0, h->p, nullptr, h->aa_cols, h->scal, h->aa_x, h->aa_c, h->p);
        // one outer iteration as a HIP graph when it needs no host decisions
        const bool use_graph = h->prm.restol_pson <= 0.0 && !h->timing && (h->prm.launch_mode & 1) == 0 &&
                               (h->prm.itermax_gc <= 0 || (h->gsched_ok && h->prm.mix_GC_form != 1));
        hipGraphExec_t &gexec = own.gexec;
        int vc64 = 0, vc32 = 0, gsw = 0;
        if (use_graph) {
            HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
            int brc = aa_body(h, G, &vc64, &vc32, &gsw);
            hipError_t ce = hipStreamEndCapture(h->st, &own.graph);
            if (brc) return brc;
            if (ce != hipSuccess) return fail(-3, std::string("graph capture failed: ") + hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&gexec, own.graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) return fail(-3, std::string("graph instantiate failed: ") + hipGetErrorString(ie));
        }
        while (iter < h->prm.itermax_pson) {
            if (gexec) {
                HIPCHK(hipGraphLaunch(gexec, h->st));
            } else {
                vc64 = vc32 = gsw = 0;
                if ((rc = aa_body(h, G, &vc64, &vc32, &gsw))) return rc;
            }
            st.vcycles_fp64 += vc64;
            st.vcycles_fp32 += vc32;
            st.gcm_sweeps += gsw;
            if (mixed || fold) st.refine_steps += h->prm.iterInter + 1;
            iter++;
            if ((rc = read_scalars(h, 4))) return rc;
            double s2 = h->hscal[0], d2 = h->hscal[1], maxf = h->hscal[2], maxg = h->hscal[3];
            r3[0] = std::sqrt(s2 / node_norm(h));
            r3[1] = maxf;
            r3[2] = (d2 > 0.0) ? std::sqrt(s2 / d2) : std::sqrt(s2);
            st.gcm_change = maxf;
            if (h->prm.infoconv == 1) printf("solve_poisson: %5d  rel %12.5E  max|f| %12.5E\n", iter, r3[2], maxf);
            if (r3[2] <= h->prm.rel_tol && maxf <= h->prm.rel_tol * std::max(maxg, 1e-300)) { converged = 1; break; }
            if (iter >= h->prm.itermax_pson) break;
            if (E64) aa_next<double>(h->st, d0, G, E64, h->aa_cols, h->scal, h->aa_x, h->aa_c, h->p);
            else aa_next<float>(h->st, d0, G, E, h->aa_cols, h->scal, h->aa_x, h->aa_c, h->p);
        }
        // the answer is the last map evaluation G (with the fold: the last
        // iterate after its GCM update, whose residual the stop test measured)
        if (fold && h->prm.itermax_gc > 0 && iter > 0)
            gcm_fold<float>(h->st, h->gnode, h->glevel_off.back(), h->gdelta, nullptr, h->p, 0);
        if (mixed && !fold) {
            add_into(h->st, d0, h->aa_g, nullptr, h->p);
            bc_fill<double>(h->st, d0, h->p, bcinfo(h, 0, false));
        }
    }
    hipEventRecord(e1, h->st);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    st.outer_iters = iter;
    st.maxres = r3[0];
    st.maxres2 = r3[1];
    st.rel_resid = r3[2];
    st.converged = converged;
    st.time_ms = ms;
    if (st_out) *st_out = st;
    return st.converged ? 0 : 1;
}

int check_handle(const fc_pson *h)
{
    if (!h) return fail(-1, "null handle");
    return 0;
}

}  // namespace

// ===========================================================================
extern "C" {

const char *fc_pson_version(void) { return "fc_poisson gfx950 0.1"; }
const char *fc_pson_last_error(void) { return g_err.c_str(); }

void fc_pson_default_params(fc_pson_params *p)
{
    memset(p, 0, sizeof(*p));
    p->it_solver_type = FC_IT_SOLVER_TYPE_MG;
    p->omega_pson = 1.0;
    p->itermax_pson = 200;
    p->restol_pson = 1e-10;
    p->iterInter = 1;
    p->restol_gcp = 1e-10;
    p->itermax_gc = 10;
    p->mix_GC_form = 0;
    p->is_PBC_homogeneous = 1;
    for (int f = 0; f < 6; ++f) { p->pbc[f] = FC_PBC_NEUMANN; p->ppp[f] = 0.0; p->bc_vel[f] = FC_BC_TYPE_DIRICHLET; }
    p->precision = FC_PREC_FP64;
    p->outer_rule = FC_OUTER_REFERENCE;
    p->rel_tol = 1e-8;
    p->fp32_cycles = 1;
    p->vc_walls = 1;
    p->coarse_solve = 0;
}

static int create_impl(const fc_grid *gr, const fc_pson_params *p, Arena *pool, hipStream_t stream, fc_pson **out)
{
    if (!gr || !p || !out) return fail(-1, "null argument");
    if (gr->nx < 4 || gr->ny < 4 || gr->nz < 4) return fail(-1, "grid needs at least 4 nodes per axis");
    if (gr->nx > 1025 || gr->ny > 1025 || gr->nz > 1025)
        return fail(-1, "at most 1025 nodes per axis (1024 cells per line in the line solver)");
    if (!gr->x || !gr->y || !gr->z) return fail(-1, "null grid coordinates");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-5, "no HIP device available");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail(-5, "no current HIP device");
    {
        const int rc0 = resolve_scratch(dev);
        if (rc0) return rc0;
    }
    fc_pson *h = new fc_pson();
    h->device = dev;
    h->prm = *p;
    h->nx = gr->nx; h->ny = gr->ny; h->nz = gr->nz;
    h->arena = pool;
    if (stream) {
        h->st = stream;
        h->own_stream = false;
    } else if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(-3, "hipStreamCreate failed");
    }
    // level-1 metrics
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    HostMet l1;
    l1.x.assign(gr->x, gr->x + nx + 2);
    l1.y.assign(gr->y, gr->y + ny + 2);
    l1.z.assign(gr->z, gr->z + nz + 2);
    grid_metrics(nx, l1.x, l1.xc, h->dx1, l1.dxinv, h->dxc1, l1.dxcinv);
    grid_metrics(ny, l1.y, l1.yc, h->dy1, l1.dyinv, h->dyc1, l1.dycinv);
    grid_metrics(nz, l1.z, l1.zc, h->dz1, l1.dzinv, h->dzc1, l1.dzcinv);
    // MG_initial copies xc/dxinv/dxcinv for 0..nx (CORE_MG.F90:202-218)
    l1.xc[nx + 1] = 0; l1.dxinv[nx + 1] = 0; l1.dxcinv[nx + 1] = 0;
    l1.yc[ny + 1] = 0; l1.dyinv[ny + 1] = 0; l1.dycinv[ny + 1] = 0;
    l1.zc[nz + 1] = 0; l1.dzinv[nz + 1] = 0; l1.dzcinv[nz + 1] = 0;
    int mgx = auto_levels(nx, p->mgrids_x), mgy = auto_levels(ny, p->mgrids_y), mgz = auto_levels(nz, p->mgrids_z);
    int ml = std::max(mgx, std::max(mgy, mgz));
    h->nlev = ml;
    h->hm.push_back(l1);
    std::vector<int> I(ml), J(ml), K(ml);
    I[0] = nx; J[0] = ny; K[0] = nz;
    for (int n = 1; n < ml; ++n) {
        I[n] = (n + 1 <= mgx) ? coarsen(I[n - 1]) : I[n - 1];
        J[n] = (n + 1 <= mgy) ? coarsen(J[n - 1]) : J[n - 1];
        K[n] = (n + 1 <= mgz) ? coarsen(K[n - 1]) : K[n - 1];
        HostMet c;
        coarse_axis(I[n], I[n - 1], h->hm[n - 1].x, l1.x[1], l1.x[nx], c.x, c.xc, c.dxinv, c.dxcinv);
        coarse_axis(J[n], J[n - 1], h->hm[n - 1].y, l1.y[1], l1.y[ny], c.y, c.yc, c.dyinv, c.dycinv);
        coarse_axis(K[n], K[n - 1], h->hm[n - 1].z, l1.z[1], l1.z[nz], c.z, c.zc, c.dzinv, c.dzcinv);
        h->hm.push_back(c);
    }
    for (int n = 0; n < ml; ++n) {
        LevelDims d;
        d.nx = I[n]; d.ny = J[n]; d.nz = K[n];
        d.px = (int)pad16(d.nx + 2);
        d.sy = d.px;
        d.sz = (long)d.px * (d.ny + 2);
        d.n = d.sz * (d.nz + 2);
        d.ic = (n > 0 && I[n] != I[n - 1]) ? 1 : 0;
        d.jc = (n > 0 && J[n] != J[n - 1]) ? 1 : 0;
        d.kc = (n > 0 && K[n] != K[n - 1]) ? 1 : 0;
        h->dims.push_back(d);
    }
    // the line sweeps address a field with 32-bit byte offsets
    if (h->dims[0].n * (long)sizeof(double) > 2147483647L) {
        fc_pson_destroy(h);
        return fail(-1, "grid too large: one FP64 field must stay below 2 GiB");
    }
    int rc = 0;
    h->m64.resize(ml); h->m32.resize(ml); h->L64.resize(ml); h->ibk.resize(ml); h->igk.resize(ml);
    h->rflag.assign(ml, nullptr); h->pflag.assign(ml, nullptr);
    h->wmask.assign(ml, nullptr); h->wpflag.assign(ml, nullptr);
    for (int n = 0; n < ml && !rc; ++n) {
        rc = upload_metrics<double>(h, n, h->m64[n]);
        if (!rc) rc = upload_metrics<float>(h, n, h->m32[n]);
        if (!rc) rc = dalloc(h, &h->ibk[n], h->dims[n].n);
        if (!rc) rc = dalloc(h, &h->igk[n], h->dims[n].n);
        if (!rc && n > 0) rc = dalloc(h, &h->rflag[n], h->dims[n].n);
        if (!rc && n > 0) rc = dalloc(h, &h->pflag[n], h->dims[n].n);
        if (!rc) rc = dalloc(h, &h->wmask[n], h->dims[n].n);
        if (!rc && n > 0) rc = dalloc(h, &h->wpflag[n], h->dims[n].n);
        if (!rc) rc = dalloc(h, &h->L64[n].res, h->dims[n].n);
        if (!rc && n > 0) rc = dalloc(h, &h->L64[n].phi, h->dims[n].n);
        if (!rc && n > 0) rc = dalloc(h, &h->L64[n].rhs, h->dims[n].n);
    }
    const LevelDims &d0 = h->dims[0];
    if (!rc) rc = dalloc(h, &h->p, d0.n);
    if (!rc) rc = dalloc(h, &h->div, d0.n);
    if (!rc) rc = dalloc(h, &h->pprev, d0.n);
    if (!rc) rc = dalloc(h, &h->ptemp, d0.n);
    if (!rc) {
        h->L64[0].phi = h->p;
        h->L64[0].rhs = h->div;
    }
    for (int f = 0; f < 6 && !rc; ++f) {
        size_t n = (f < 2) ? (size_t)(ny + 2) * (nz + 2) : (f < 4) ? (size_t)(nx + 2) * (nz + 2) : (size_t)(nx + 2) * (ny + 2);
        rc = dalloc(h, &h->pg[f], n);
    }
    // level-1 1-D arrays for the GCM and rhs kernels
    if (!rc) {
        std::vector<double> v;
        auto app = [&](const std::vector<double> &a) { v.insert(v.end(), a.begin(), a.end()); };
        app(l1.dxcinv); app(l1.dycinv); app(l1.dzcinv); app(l1.dxinv); app(l1.dyinv); app(l1.dzinv);
        app(h->dx1); app(h->dy1); app(h->dz1);
        rc = dalloc(h, &h->d_l1w, v.size());
        if (!rc) {
            hipMemcpyAsync(h->d_l1w, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, h->st);
            const double *b = h->d_l1w;
            size_t ox = 0, oy = nx + 2, oz = oy + ny + 2;
            size_t s = (size_t)(nx + ny + nz + 6);
            h->l1met[0] = b + ox; h->l1met[1] = b + oy; h->l1met[2] = b + oz;
            h->l1met[3] = b + s + ox; h->l1met[4] = b + s + oy; h->l1met[5] = b + s + oz;
            h->l1rhs[0] = h->l1met[3]; h->l1rhs[1] = h->l1met[4]; h->l1rhs[2] = h->l1met[5];
            h->l1rhs[3] = b + 2 * s + ox; h->l1rhs[4] = b + 2 * s + oy; h->l1rhs[5] = b + 2 * s + oz;
        }
    }
    if (!rc) {
        long np = std::max<long>((long)(d0.ny + 2) * (d0.nz + 2), 8192);
        np = std::max<long>(np, (long)(d0.n / 256 + 1024));
        // row-strided reductions store ROW_COMPONENTS x ROW_BLOCKS partials
        np = std::max<long>(np, (long)ROW_COMPONENTS * ROW_BLOCKS / 2);
        h->npartial = np;
        rc = dalloc(h, &h->partial, 2 * (size_t)np);
    }
    if (!rc) rc = dalloc(h, &h->scal, 128);
    if (!rc) {
        // pinned for the per-iteration read-back; pooled handles read back
        // through their server and keep a pageable mirror
        if (h->arena) {
            h->hscal = (double *)std::calloc(128, sizeof(double));
            if (!h->hscal) rc = fail(-4, "out of host memory");
        } else if (hipHostMalloc((void **)&h->hscal, 128 * sizeof(double), 0) != hipSuccess) {
            rc = fail(-4, "hipHostMalloc failed");
        }
    }
    if (!rc && hipStreamSynchronize(h->st) != hipSuccess) rc = fail(-3, "stream sync failed after allocation");
    if (rc) {
        fc_pson_destroy(h);
        return rc;
    }
    *out = h;
    return 0;
}

int fc_pson_create(const fc_grid *gr, const fc_pson_params *p, fc_pson **out)
{
    return create_impl(gr, p, nullptr, nullptr, out);
}

int fc_pson_set_params(fc_pson *h, const fc_pson_params *p)
{
    if (check_handle(h) || !p) return fail(-1, "null argument");
    h->prm = *p;
    drop_coarse_direct(h);
    return 0;
}

int fc_pson_destroy(fc_pson *h)
{
    if (!h) return 0;
    if (h->st) hipStreamSynchronize(h->st);
    for (auto &m : h->m64) dfree(h, m.buf);
    for (auto &m : h->m32) dfree(h, m.buf);
    for (size_t n = 0; n < h->L64.size(); ++n) {
        if (n > 0) { dfree(h, h->L64[n].phi); dfree(h, h->L64[n].rhs); }
        dfree(h, h->L64[n].res);
    }
    for (auto &L : h->L32) { dfree(h, L.phi); dfree(h, L.rhs); dfree(h, L.res); }
    if (!h->L64c.empty()) { dfree(h, h->L64c[0].phi); dfree(h, h->L64c[0].rhs); }
    for (auto *q : h->ibk) dfree(h, q);
    for (auto *q : h->igk) dfree(h, q);
    for (auto *q : h->rflag) dfree(h, q);
    for (auto *q : h->pflag) dfree(h, q);
    for (auto *q : h->wmask) dfree(h, q);
    for (auto *q : h->wpflag) dfree(h, q);
    dfree(h, h->cd_M);
    dfree(h, h->cd_q);
    for (double *q : {h->p, h->div, h->pprev, h->ptemp, h->d_l1w, h->partial, h->scal}) dfree(h, q);
    for (int f = 0; f < 6; ++f) dfree(h, h->pg[f]);
    dfree(h, h->gbuf);
    dfree(h, h->d_glevel);
    for (double *q : {h->aa_x, h->aa_g}) dfree(h, q);
    for (float *q : {h->aa_fprev, h->aa_c}) dfree(h, q);
    for (int i = 0; i < 8; ++i) {
        dfree(h, h->aa_cols.dF[i]);
        dfree(h, h->aa_cols.dG[i]);
    }
    dfree(h, h->aa_ring);
    for (double *q : {h->vel, h->velbc, h->projmet, h->ad_coefs, h->ad_rhs, h->ad_partial, h->d_ynode, h->gv_buf,
                      h->gv_vals, h->visc_tot})
        dfree(h, q);
    dfree(h, h->upw);
    dfree(h, h->pzero);
    dfree(h, h->gvel);
    for (hipEvent_t e : h->ev) hipEventDestroy(e);
    if (h->hscal) {
        if (h->arena) std::free(h->hscal);
        else hipHostFree(h->hscal);
    }
    if (h->st && h->own_stream) hipStreamDestroy(h->st);
    delete h;
    return 0;
}

int fc_pson_levels(const fc_pson *h, int *nlev, int *nxl, int *nyl, int *nzl)
{
    if (!h) return fail(-1, "null handle");
    if (nlev) *nlev = h->nlev;
    for (int n = 0; n < h->nlev; ++n) {
        if (nxl) nxl[n] = h->dims[n].nx;
        if (nyl) nyl[n] = h->dims[n].ny;
        if (nzl) nzl[n] = h->dims[n].nz;
    }
    return 0;
}

int fc_pson_set_masks(fc_pson *h, const int8_t *iblank, const int8_t *ghost, const int8_t *dead)
{
    if (check_handle(h) || !iblank || !ghost) return fail(-1, "null argument");
    size_t n = (size_t)(h->nx + 2) * (h->ny + 2) * (h->nz + 2);
    h->h_iblank.assign(iblank, iblank + n);
    h->h_ghost.assign(ghost, ghost + n);
    if (dead) h->h_dead.assign(dead, dead + n); else h->h_dead.assign(n, 0);
    std::atomic<int> bad(0);
    const long plane = (long)(h->nx + 2) * (h->ny + 2);
    par_planes(0, h->nz + 1, plane, [&](int k) {
        const int8_t *ib = h->h_iblank.data() + (size_t)k * plane, *gh = h->h_ghost.data() + (size_t)k * plane;
        int b = 0;
        for (long q = 0; q < plane; ++q) b |= (ib[q] & ~1) | (gh[q] & ~1);
        if (b) bad.store(1);
    });
    if (bad.load()) return fail(-1, "iblank and ghostCellMark must be 0 or 1");
    h->have_masks = true;
    h->gcm_dirty = true;
    h->pzero_dirty = true;
    drop_coarse_direct(h);
    return prepare_masks(h);
}

int fc_pson_set_gcm(fc_pson *h, int nGhost, const int *iGhost, const int *jGhost, const int *kGhost,
                    const int *iCellIndex, const int *jCellIndex, const int *kCellIndex, const double *coeffGCMN,
                    const double *alphaGhost, const double *dpdnBodyIntercept, int nDead, const int *iDead,
                    const int *jDead, const int *kDead, const int *iDeadCellIndex, const int *jDeadCellIndex,
                    const int *kDeadCellIndex, const double *coeffGCMDeadN, const double *probeLengthDead,
                    const double *dpdnInterceptDead, const double *xyzDCIP)
{
    if (check_handle(h)) return -1;
    if (nGhost < 0 || nDead < 0) return fail(-1, "negative list length");
    if (nGhost > 0 && (!iGhost || !jGhost || !kGhost || !iCellIndex || !jCellIndex || !kCellIndex || !coeffGCMN))
        return fail(-1, "null ghost-cell list");
    if (nDead > 0 && (!iDead || !jDead || !kDead || !iDeadCellIndex || !jDeadCellIndex || !kDeadCellIndex ||
                      !coeffGCMDeadN || !probeLengthDead || !xyzDCIP))
        return fail(-1, "null dead-cell list");
    auto &g = h->hg;
    g.iG.assign(iGhost, iGhost + nGhost); g.jG.assign(jGhost, jGhost + nGhost); g.kG.assign(kGhost, kGhost + nGhost);
    g.iC.assign(iCellIndex, iCellIndex + nGhost); g.jC.assign(jCellIndex, jCellIndex + nGhost);
    g.kC.assign(kCellIndex, kCellIndex + nGhost);
    g.cN.assign(coeffGCMN, coeffGCMN + 8 * (size_t)nGhost);
    g.alpha.assign(nGhost, 0.0);
    if (alphaGhost) g.alpha.assign(alphaGhost, alphaGhost + nGhost);
    g.dpdnG.assign(nGhost, 0.0);
    if (dpdnBodyIntercept) g.dpdnG.assign(dpdnBodyIntercept, dpdnBodyIntercept + nGhost);
    g.iD.assign(iDead, iDead + nDead); g.jD.assign(jDead, jDead + nDead); g.kD.assign(kDead, kDead + nDead);
    g.iDC.assign(iDeadCellIndex, iDeadCellIndex + nDead); g.jDC.assign(jDeadCellIndex, jDeadCellIndex + nDead);
    g.kDC.assign(kDeadCellIndex, kDeadCellIndex + nDead);
    g.cDN.assign(coeffGCMDeadN, coeffGCMDeadN + 8 * (size_t)nDead);
    g.plD.assign(probeLengthDead, probeLengthDead + nDead);
    g.dpdnD.assign(nDead, 0.0);
    if (dpdnInterceptDead) g.dpdnD.assign(dpdnInterceptDead, dpdnInterceptDead + nDead);
    g.ipD.assign(xyzDCIP, xyzDCIP + 3 * (size_t)nDead);
    h->gcm_dirty = true;
    h->gvel_dirty = true;
    return 0;
}

int fc_pson_set_bc(fc_pson *h, const double *pgx1, const double *pgx2, const double *pgy1, const double *pgy2,
                   const double *pgz1, const double *pgz2, double outflow_area)
{
    if (check_handle(h)) return -1;
    const double *src[6] = {pgx1, pgx2, pgy1, pgy2, pgz1, pgz2};
    int nx = h->nx, ny = h->ny, nz = h->nz;
    for (int f = 0; f < 6; ++f) {
        size_t n = (f < 2) ? (size_t)(ny + 2) * (nz + 2) : (f < 4) ? (size_t)(nx + 2) * (nz + 2) : (size_t)(nx + 2) * (ny + 2);
        if (src[f]) HIPCHK(hipMemcpyAsync(h->pg[f], src[f], sizeof(double) * n, hipMemcpyHostToDevice, h->st));
        else HIPCHK(hipMemsetAsync(h->pg[f], 0, sizeof(double) * n, h->st));
    }
    h->outflow_area = outflow_area;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_pson_get_bc(fc_pson *h, double *pgx1, double *pgx2, double *pgy1, double *pgy2, double *pgz1, double *pgz2)
{
    if (check_handle(h)) return -1;
    double *dst[6] = {pgx1, pgx2, pgy1, pgy2, pgz1, pgz2};
    int nx = h->nx, ny = h->ny, nz = h->nz;
    for (int f = 0; f < 6; ++f) {
        size_t n = (f < 2) ? (size_t)(ny + 2) * (nz + 2) : (f < 4) ? (size_t)(nx + 2) * (nz + 2) : (size_t)(nx + 2) * (ny + 2);
        if (dst[f]) HIPCHK(hipMemcpyAsync(dst[f], h->pg[f], sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_pson_upload(fc_pson *h, const double *pPrime, const double *div, const double *p_prev)
{
    if (check_handle(h)) return -1;
    int rc;
    if (pPrime && (rc = to_device(h, h->p, pPrime))) return rc;
    if (div && (rc = to_device(h, h->div, div))) return rc;
    if (p_prev) { if ((rc = to_device(h, h->pprev, p_prev))) return rc; }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_pson_download(fc_pson *h, double *pPrime, double *res_laplacian)
{
    if (check_handle(h)) return -1;
    int rc;
    if (pPrime && (rc = to_host(h, pPrime, h->p))) return rc;
    if (res_laplacian && h->g.nG > 0 && h->reslap)
        HIPCHK(hipMemcpyAsync(res_laplacian, h->reslap, sizeof(double) * h->g.nG, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_pson_solve_resident(fc_pson *h, fc_pson_stats *st)
{
    if (check_handle(h)) return -1;
    return solve_resident(h, st);
}

int fc_pson_solve(fc_pson *h, double *pPrime, const double *div, const double *p_prev, double *res_laplacian,
                  fc_pson_stats *st)
{
    if (check_handle(h) || !pPrime || !div) return fail(-1, "null argument");
    int rc = fc_pson_upload(h, pPrime, div, p_prev);
    if (rc) return rc;
    if (!p_prev) { HIPCHK(hipMemsetAsync(h->pprev, 0, sizeof(double) * h->dims[0].n, h->st)); }
    int src = solve_resident(h, st);
    if (src < 0) return src;
    rc = fc_pson_download(h, pPrime, res_laplacian);
    if (rc) return rc;
    return src;
}

int fc_pson_mg_solve(fc_pson *h, double *var, const double *rr, fc_pson_stats *st)
{
    if (check_handle(h) || !var || !rr) return fail(-1, "null argument");
    int rc = fc_pson_upload(h, var, rr, nullptr);
    if (rc) return rc;
    if (h->prm.precision == FC_PREC_MIXED && (rc = ensure_mixed(h))) return rc;
    fc_pson_stats loc;
    memset(&loc, 0, sizeof(loc));
    double resm = 0.0;
    if ((rc = mg_solve_resident(h, &loc, &resm))) return rc;
    loc.maxres = resm;
    if ((rc = fc_pson_download(h, var, nullptr))) return rc;
    if (st) *st = loc;
    return 0;
}

int fc_pson_itsolv(fc_pson *h, double *var, const double *r)
{
    if (check_handle(h) || !var || !r) return fail(-1, "null argument");
    int rc = fc_pson_upload(h, var, r, nullptr);
    if (rc) return rc;
    itsolv_resident(h);
    return fc_pson_download(h, var, nullptr);
}

int fc_pson_ghostcell_pressure(fc_pson *h, double *pres, const double *r, const double *p_prev,
                               double *res_laplacian, fc_pson_stats *st)
{
    if (check_handle(h) || !pres || !r) return fail(-1, "null argument");
    int rc = fc_pson_upload(h, pres, r, p_prev);
    if (rc) return rc;
    if (!p_prev) { HIPCHK(hipMemsetAsync(h->pprev, 0, sizeof(double) * h->dims[0].n, h->st)); }
    fc_pson_stats loc;
    memset(&loc, 0, sizeof(loc));
    if ((rc = gcm_update(h, &loc))) return rc;
    if (loc.gcm_change < 0.0) {
        HIPCHK(hipMemcpyAsync(h->hscal + 10, h->scal + 10, sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        loc.gcm_change = 0.0;   /* an exact pass leaves nothing to converge */
    }
    if ((rc = fc_pson_download(h, pres, res_laplacian))) return rc;
    if (st) *st = loc;
    return (loc.gcm_change > h->prm.restol_gcp) ? 1 : 0;
}

// loc (may be null): 1-based (i, j, k) of the first cell in the reference's
// (k, j, i) loop order where the largest |value| occurs; (0, 0, 0) when every
// value is zero (the reference leaves loc unset then)
static int read_loc(fc_pson *h, const double *var, const double *r, const int8_t *mask, int mode, int *loc)
{
    const LevelDims &d = h->dims[0];
    resloc(h->st, d, h->m64[0].m, var, r, mask, mode, h->partial, h->scal + 110);
    HIPCHK(hipMemcpyAsync(h->hscal + 110, h->scal + 110, 3 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    const double t = h->hscal[111];
    if (t < 0.0) { loc[0] = loc[1] = loc[2] = 0; return 0; }
    const long ti = (long)t, ni = d.nx - 1, nj = d.ny - 1;
    loc[0] = 1 + (int)(ti % ni);
    loc[1] = 1 + (int)((ti / ni) % nj);
    loc[2] = 1 + (int)(ti / (ni * nj));
    return 0;
}

int fc_pson_calc_residual(fc_pson *h, const double *var, const double *r, double *resm, int *loc)
{
    if (check_handle(h) || !var || !r || !resm) return fail(-1, "null argument");
    int rc = fc_pson_upload(h, var, r, nullptr);
    if (rc) return rc;
    resnorm<double>(h->st, h->dims[0], h->m64[0].m, h->p, h->div, h->igk[0], h->partial, h->scal);
    if ((rc = read_scalars(h, 2))) return rc;
    *resm = std::sqrt(h->hscal[0] / node_norm(h));
    if (loc) return read_loc(h, h->p, h->div, h->igk[0], 0, loc);
    return 0;
}
int fc_pson_calc_residual2(fc_pson *h, const double *var, const double *p_temp, double *resm, int *loc)
{
    if (check_handle(h) || !var || !p_temp || !resm) return fail(-1, "null argument");
    int rc;
    if ((rc = to_device(h, h->p, var))) return rc;
    if ((rc = to_device(h, h->ptemp, p_temp))) return rc;
    maxdiff(h->st, h->dims[0], h->p, h->ptemp, h->ibk[0], h->partial, h->scal + 2);
    HIPCHK(hipMemcpyAsync(h->hscal + 2, h->scal + 2, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    *resm = h->hscal[2];
    if (loc) return read_loc(h, h->p, h->ptemp, h->ibk[0], 1, loc);
    return 0;
}

// fluid area of the zero-gradient faces (enforce_global_mass_consv,
// CORE_SET_OUTER_BC.F90:584-637: (1 - iblank) of the first and last
// interior layer, summed in the reference's loop order)
namespace {
int mass_consv_device(fc_pson *h, double *du, double *dv, double *dw);
}

static double outflow_area_host(const fc_pson *h, int zg)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    auto ib = [&](int i, int j, int k) {
        return (double)h->h_iblank[(size_t)i + (size_t)(nx + 2) * ((size_t)j + (size_t)(ny + 2) * k)];
    };
    double aL = 0.0, aR = 0.0, aB = 0.0, aT = 0.0, aK = 0.0, aF = 0.0;
    for (int k = 1; k <= nz - 1; ++k)
        for (int j = 1; j <= ny - 1; ++j) {
            aL = aL + h->dy1[j] * h->dz1[k] * (1.0 - ib(1, j, k));
            aR = aR + h->dy1[j] * h->dz1[k] * (1.0 - ib(nx - 1, j, k));
        }
    for (int k = 1; k <= nz - 1; ++k)
        for (int i = 1; i <= nx - 1; ++i) {
            aB = aB + h->dx1[i] * h->dz1[k] * (1.0 - ib(i, 1, k));
            aT = aT + h->dx1[i] * h->dz1[k] * (1.0 - ib(i, ny - 1, k));
        }
    for (int j = 1; j <= ny - 1; ++j)
        for (int i = 1; i <= nx - 1; ++i) {
            aK = aK + h->dx1[i] * h->dy1[j] * (1.0 - ib(i, j, 1));
            aF = aF + h->dx1[i] * h->dy1[j] * (1.0 - ib(i, j, nz - 1));
        }
    double area = 0.0;
    if (zg & 1) area = area + aL;
    if (zg & 2) area = area + aR;
    if (zg & 4) area = area + aB;
    if (zg & 8) area = area + aT;
    if (zg & 16) area = area + aK;
    if (zg & 32) area = area + aF;
    return area;
}

int fc_pson_rhs(fc_pson *h, double *fu, double *fv, double *fw, double dtinv, double *div, double *sumd,
                double *resDiv)
{
    if (check_handle(h) || !fu || !fv || !fw || !div) return fail(-1, "null argument");
    if (!h->have_masks) return fail(-1, "fc_pson_set_masks must be called before rhs_poisson");
    const LevelDims &d = h->dims[0];
    const int zg = zg_mask(h);
    const bool mix = h->prm.mix_GC_form == 1;
    if (mix) {
        int rc0 = finalize_gcm(h);
        if (!rc0 && h->gvel_dirty) rc0 = prepare_gvel(h);
        if (rc0) return rc0;
    }
    double *buf = nullptr;
    int rc = dalloc(h, &buf, 3 * (size_t)d.n);
    if (rc) return rc;
    h->dev_bytes -= sizeof(double) * 3 * (size_t)d.n;
    double *du = buf, *dv = buf + d.n, *dw = buf + 2 * d.n;
    // the scratch buffer is released on every exit path
    auto body = [&]() -> int {
        int r;
        if ((r = to_device(h, du, fu)) || (r = to_device(h, dv, fv)) || (r = to_device(h, dw, fw))) return r;
        // enforce_global_mass_consv (CORE_SET_OUTER_BC.F90:515-735)
        if ((r = mass_consv_device(h, du, dv, dw))) return r;
        fill<double>(h->st, d.n, h->div, 0.0);
        rhs_poisson(h->st, d, du, dv, dw, h->igk[0], h->l1rhs, dtinv, h->div, h->partial, h->scal);
        if (mix) ghost_massflux(h->st, d, h->gvel, h->ngvel, du, dv, dw, h->l1rhs, dtinv, h->div, h->partial,
                                h->scal + 118);
        HIPCHK(hipMemcpyAsync(h->hscal, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
        if (mix) HIPCHK(hipMemcpyAsync(h->hscal + 118, h->scal + 118, sizeof(double), hipMemcpyDeviceToHost, h->st));
        if ((r = to_host(h, div, h->div))) return r;
        if (zg && ((r = to_host(h, fu, du)) || (r = to_host(h, fv, dv)) || (r = to_host(h, fw, dw)))) return r;
        HIPCHK(hipStreamSynchronize(h->st));
        if (sumd) *sumd = h->hscal[0] / dtinv + (mix ? h->hscal[118] : 0.0);
        if (resDiv) *resDiv = h->hscal[1];
        return 0;
    };
    rc = body();
    hipStreamSynchronize(h->st);
    dfree(h, buf);
    return rc;
}

int fc_pson_get_outflow(fc_pson *h, double *outflow_area, double *correction_vel)
{
    if (check_handle(h)) return -1;
    if (outflow_area) *outflow_area = h->outflow_area;
    if (correction_vel) *correction_vel = h->outflow_corr;
    return 0;
}

int fc_pson_rhs_noalloc(fc_pson *h, const double *face_u, const double *face_v, const double *face_w, double dtinv,
                        double *div, double *sums, void *stream)
{
    if (check_handle(h) || !face_u || !face_v || !face_w || !div || !sums) return fail(-1, "null argument");
    for (int f = 0; f < 6; ++f)
        if (h->prm.bc_vel[f] == FC_BC_TYPE_ZERO_GRADIENT)
            return fail(-2, "rhs_poisson with zero-gradient outflow faces is not supported by this build");
    if (h->prm.mix_GC_form == 1) return fail(-2, "rhs_poisson with mix_GC_form=1 is not supported by this build");
    const LevelDims &dm = h->dims[0];
    const long rows = (long)(dm.ny - 1) * (dm.nz - 1);
    if (rows > FC_NOALLOC_MAX_ROWS)
        return fail(-1, "fc_pson_rhs_noalloc takes at most FC_NOALLOC_MAX_ROWS = " +
                            std::to_string(FC_NOALLOC_MAX_ROWS) + " interior (j,k) rows; this grid has " +
                            std::to_string(rows));
    double *scr = scratch_of(h->device);
    if (!scr) return fail(-3, "the scratch address of this handle's device was not resolved");
    // the caller's fields in the unpadded Fortran layout (0:nx+1, 0:ny+1, 0:nz+1)
    LevelDims df = dm;
    df.px = dm.nx + 2;
    df.sy = df.px;
    df.sz = (long)df.px * (dm.ny + 2);
    df.n = df.sz * (dm.nz + 2);
    hipStream_t st = stream ? (hipStream_t)stream : h->st;
    rhs_poisson2(st, df, dm, face_u, face_v, face_w, h->igk[0], h->l1rhs, dtinv, div, scr, sums);
    rhs_sumd(st, sums, dtinv);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(-3, std::string("HIP error: ") + hipGetErrorString(e));
    return 0;
}

int fc_tdma_batched(int nsys, int n, const double *a, const double *b, const double *c, const double *r, double *u,
                    int prec)
{
    if (nsys <= 0 || n <= 0 || !a || !b || !c || !r || !u) return fail(-1, "bad arguments");
    if (n > 576) return fail(-1, "fc_tdma_batched supports at most 576 rows per system");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-5, "no HIP device available");
    size_t bytes = sizeof(double) * (size_t)nsys * n;
    double *d = nullptr;
    HIPCHK(hipMalloc(&d, 5 * bytes));
    size_t N = (size_t)nsys * n;
    hipMemcpy(d, a, bytes, hipMemcpyHostToDevice);
    hipMemcpy(d + N, b, bytes, hipMemcpyHostToDevice);
    hipMemcpy(d + 2 * N, c, bytes, hipMemcpyHostToDevice);
    hipMemcpy(d + 3 * N, r, bytes, hipMemcpyHostToDevice);
    int rc = tdma_batched(0, nsys, n, d, d + N, d + 2 * N, d + 3 * N, d + 4 * N, prec);
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess && rc == 0) e = hipMemcpy(u, d + 4 * N, bytes, hipMemcpyDeviceToHost);
    hipFree(d);
    if (rc) return fail(rc, "unsupported system size");
    if (e != hipSuccess) return fail(-3, std::string("HIP error: ") + hipGetErrorString(e));
    return 0;
}

size_t fc_pson_device_bytes(const fc_pson *h) { return h ? h->dev_bytes : 0; }

int fc_pson_reset(fc_pson *h)
{
    if (check_handle(h)) return -1;
    fill<double>(h->st, h->dims[0].n, h->p, 0.0);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

// ---------------------------------------------------------------------------
// Velocity projection (the calls after solve_poisson, CORE_TIME_STEP.F90:
// 211-222).

namespace {

// update_pressure's zero set at one rank (CORE_FRAC_STEP_ROUTINES.F90:
// 212-297): solid cells that are not dead-cell candidates over i = 0..nx,
// j = 0..ny+1, k = 0..nz+1 (:225-236), and the domain-edge lines (:250-294).
// The k = zc_start-2 = -1 plane named at :279 and :282 lies outside the
// array at one rank and is left out.
int prepare_pzero(fc_pson *h)
{
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const size_t hx = nx + 2, hy = ny + 2;
    std::vector<int8_t> z(hx * hy * (size_t)(nz + 2), 0);
    auto at = [&](int i, int j, int k) -> int8_t & { return z[(size_t)i + hx * ((size_t)j + hy * k)]; };
    for (int k = 0; k <= nz + 1; ++k)
        for (int j = 0; j <= ny + 1; ++j)
            for (int i = 0; i <= nx; ++i) {
                const size_t q = (size_t)i + hx * ((size_t)j + hy * k);
                if (h->h_iblank[q] != 0 && h->h_dead[q] != -1) at(i, j, k) = 1;
            }
    for (int k : {0, nz, nz + 1}) {
        for (int j : {0, ny})
            for (int i = 0; i <= nx; ++i) at(i, j, k) = 1;
        for (int i : {0, nx})
            for (int j = 0; j <= ny + 1; ++j) at(i, j, k) = 1;
    }
    for (int i : {0, nx})
        for (int j : {0, ny})
            for (int k = 0; k <= nz + 1; ++k) at(i, j, k) = 1;
    int rc = 0;
    if (!h->pzero && (rc = dalloc(h, &h->pzero, h->dims[0].n))) return rc;
    if ((rc = to_device(h, h->pzero, z.data()))) return rc;
    h->pzero_dirty = false;
    return 0;
}

// the ghost list of correct_ghostcell_velocity; every cell must be interior
// and listed once (the kernel updates the cells in parallel)
int prepare_gvel(fc_pson *h)
{
    const auto &g = h->hg;
    const int n = (int)g.iG.size();
    std::vector<GhostVel> r(n);
    std::vector<uint8_t> seen((size_t)(h->nx + 2) * (h->ny + 2) * (h->nz + 2), 0);
    for (int t = 0; t < n; ++t) {
        const int i = g.iG[t], j = g.jG[t], k = g.kG[t];
        if (i < 1 || i > h->nx - 1 || j < 1 || j > h->ny - 1 || k < 1 || k > h->nz - 1)
            return fail(-1, "correct_ghostcell_velocity: ghost cell outside the interior");
        const size_t q = (size_t)i + (size_t)(h->nx + 2) * ((size_t)j + (size_t)(h->ny + 2) * k);
        if (seen[q]) return fail(-1, "correct_ghostcell_velocity: ghost cell listed twice");
        seen[q] = 1;
        r[t] = GhostVel{i, j, k, 0, g.alpha[t]};
    }
    dfree(h, h->gvel);
    h->gvel = nullptr;
    h->ngvel = n;
    int rc = 0;
    if (n > 0) {
        if ((rc = dalloc(h, &h->gvel, (size_t)n))) return rc;
        HIPCHK(hipMemcpyAsync(h->gvel, r.data(), sizeof(GhostVel) * n, hipMemcpyHostToDevice, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
    }
    h->gvel_dirty = false;
    return 0;
}

size_t velbc_len(const fc_pson *h, int a)
{
    const size_t hx = h->nx + 2, hy = h->ny + 2, hz = h->nz + 2;
    return a == 0 ? 2 * hy * hz : a == 1 ? hx * 2 * hz : hx * hy * 2;
}

// the outer-face array of axis a (x, y, z) and component c (u, v, w); the
// face-normal ones (bcxu, bcyv, bczw) are those of face_vel
double *vbc_ptr(fc_pson *h, int a, int c)
{
    const size_t lx = velbc_len(h, 0), ly = velbc_len(h, 1), lz = velbc_len(h, 2);
    const size_t off = a == 0 ? c * lx : a == 1 ? 3 * lx + c * ly : 3 * lx + 3 * ly + c * lz;
    return h->velbc + off;
}

// buffers and metrics of the projection, allocated at the first use
int ensure_proj(fc_pson *h)
{
    if (!h->have_masks) return fail(-1, "projection before fc_pson_set_masks");
    int rc = 0;
    if (!h->vel) {
        const LevelDims &d = h->dims[0];
        if ((rc = dalloc(h, &h->vel, 6 * (size_t)d.n))) return rc;
        fill<double>(h->st, 6 * d.n, h->vel, 0.0);
        // the nine outer-face arrays: bcxu, bcxv, bcxw, bcyu, ..., bczw
        const size_t nb = 3 * (velbc_len(h, 0) + velbc_len(h, 1) + velbc_len(h, 2));
        if ((rc = dalloc(h, &h->velbc, nb))) return rc;
        fill<double>(h->st, (long)nb, h->velbc, 0.0);
        if ((rc = dalloc(h, &h->d_ynode, (size_t)h->ny + 2))) return rc;
        HIPCHK(hipMemcpyAsync(h->d_ynode, h->hm[0].y.data(), sizeof(double) * (h->ny + 2), hipMemcpyHostToDevice,
                              h->st));
        // fx, fy, fz (CORE_GRID.F90:184-205) and dxinv, dyinv, dzinv
        // (CORE_GRID.F90:120-139), each zero-padded
        const int n3[3] = {h->nx, h->ny, h->nz};
        const HostMet &m = h->hm[0];
        const std::vector<double> *xs[3] = {&m.x, &m.y, &m.z};
        size_t span[3], tot = 0;
        for (int a = 0; a < 3; ++a) { span[a] = (size_t)pad16(n3[a] + 2) + 16; tot += 2 * span[a]; }
        std::vector<double> hv(tot, 0.0);
        size_t off = 0, offs[6];
        for (int a = 0; a < 3; ++a) {
            const std::vector<double> &x = *xs[a];
            const int n = n3[a];
            std::vector<double> xc(n + 2, 0.0);
            for (int i = 0; i <= n; ++i) xc[i] = 0.5 * (x[i + 1] + x[i]);
            offs[a] = off;
            for (int i = 1; i <= n; ++i) hv[off + i] = (x[i] - xc[i - 1]) / (xc[i] - xc[i - 1]);
            off += span[a];
        }
        for (int a = 0; a < 3; ++a) {
            const std::vector<double> &x = *xs[a];
            offs[3 + a] = off;
            for (int i = 0; i <= n3[a]; ++i) hv[off + i] = 1.0 / (x[i + 1] - x[i]);
            off += span[a];
        }
        if ((rc = dalloc(h, &h->projmet, tot))) return rc;
        HIPCHK(hipMemcpyAsync(h->projmet, hv.data(), sizeof(double) * tot, hipMemcpyHostToDevice, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        h->pm = ProjMet{h->projmet + offs[0], h->projmet + offs[1], h->projmet + offs[2],
                        h->projmet + offs[3], h->projmet + offs[4], h->projmet + offs[5]};
    }
    if (h->pzero_dirty && (rc = prepare_pzero(h))) return rc;
    if (h->gvel_dirty && (rc = prepare_gvel(h))) return rc;
    return 0;
}

double *vel_field(fc_pson *h, int f) { return h->vel + (size_t)f * h->dims[0].n; }

// uploads the non-null host fields into vel slots first, first+1, ...
int vel_in(fc_pson *h, int first, std::initializer_list<const double *> src)
{
    int f = first, rc = 0;
    for (const double *s : src) {
        if (s && (rc = to_device(h, vel_field(h, f), s))) return rc;
        ++f;
    }
    return 0;
}

int vel_out(fc_pson *h, int first, std::initializer_list<double *> dst)
{
    int f = first, rc = 0;
    for (double *s : dst) {
        if (s && (rc = to_host(h, s, vel_field(h, f))) ) return rc;
        ++f;
    }
    return 0;
}

void launch_correct_vel(fc_pson *h, const double *pp, double dt)
{
    correct_vel(h->st, h->dims[0], h->pm, pp, h->ibk[0], dt, vel_field(h, 0), vel_field(h, 1), vel_field(h, 2));
}

void launch_ghost_vel(fc_pson *h, const double *pv, double dt)
{
    correct_ghost_vel(h->st, h->dims[0], h->pm, h->gvel, h->ngvel, h->prm.mix_GC_form == 1 ? 1 : 0, pv, dt,
                      vel_field(h, 0), vel_field(h, 1), vel_field(h, 2));
}

void launch_face_vel(fc_pson *h)
{
    face_vel(h->st, h->dims[0], h->pm, vel_field(h, 0), vel_field(h, 1), vel_field(h, 2), vbc_ptr(h, 0, 0),
             vbc_ptr(h, 1, 1), vbc_ptr(h, 2, 2), vel_field(h, 3), vel_field(h, 4), vel_field(h, 5));
}

}  // namespace

int fc_vel_set_bc(fc_pson *h, const double *bcxu, const double *bcyv, const double *bczw)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    const double *src[3] = {bcxu, bcyv, bczw};
    for (int a = 0; a < 3; ++a) {
        const size_t n = velbc_len(h, a);
        double *b = vbc_ptr(h, a, a);
        if (src[a]) HIPCHK(hipMemcpyAsync(b, src[a], sizeof(double) * n, hipMemcpyHostToDevice, h->st));
        else fill<double>(h->st, (long)n, b, 0.0);
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_vel_get_bc(fc_pson *h, double *bcxu, double *bcyv, double *bczw)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    double *dst[3] = {bcxu, bcyv, bczw};
    for (int a = 0; a < 3; ++a) {
        const size_t n = velbc_len(h, a);
        if (dst[a]) HIPCHK(hipMemcpyAsync(dst[a], vbc_ptr(h, a, a), sizeof(double) * n, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_vel_set_outer_all(fc_pson *h, const double *const *in9)
{
    if (check_handle(h) || !in9) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) {
            const size_t n = velbc_len(h, a);
            if (in9[3 * a + c])
                HIPCHK(hipMemcpyAsync(vbc_ptr(h, a, c), in9[3 * a + c], sizeof(double) * n, hipMemcpyHostToDevice,
                                      h->st));
            else
                fill<double>(h->st, (long)n, vbc_ptr(h, a, c), 0.0);
        }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_vel_get_outer_all(fc_pson *h, double *const *out9)
{
    if (check_handle(h) || !out9) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc) return rc;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c)
            if (out9[3 * a + c])
                HIPCHK(hipMemcpyAsync(out9[3 * a + c], vbc_ptr(h, a, c), sizeof(double) * velbc_len(h, a),
                                      hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_vel_upload(fc_pson *h, const double *u, const double *v, const double *w, const double *face_u,
                  const double *face_v, const double *face_w)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    return vel_in(h, 0, {u, v, w, face_u, face_v, face_w});
}

int fc_vel_download(fc_pson *h, double *u, double *v, double *w, double *face_u, double *face_v, double *face_w)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    return vel_out(h, 0, {u, v, w, face_u, face_v, face_w});
}

int fc_correct_vel(fc_pson *h, const double *pPrime, double dt, double *u, double *v, double *w)
{
    if (check_handle(h) || !pPrime || !u || !v || !w) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc || (rc = to_device(h, h->p, pPrime)) || (rc = vel_in(h, 0, {u, v, w}))) return rc;
    launch_correct_vel(h, h->p, dt);
    HIPCHK(hipGetLastError());
    return vel_out(h, 0, {u, v, w});
}

int fc_correct_ghostcell_velocity(fc_pson *h, const double *pVar, double dt_fac, double *u, double *v, double *w)
{
    if (check_handle(h) || !pVar || !u || !v || !w) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc || (rc = to_device(h, h->p, pVar)) || (rc = vel_in(h, 0, {u, v, w}))) return rc;
    launch_ghost_vel(h, h->p, dt_fac);
    HIPCHK(hipGetLastError());
    return vel_out(h, 0, {u, v, w});
}

int fc_face_vel(fc_pson *h, const double *u, const double *v, const double *w, double *face_u, double *face_v,
                double *face_w)
{
    if (check_handle(h) || !u || !v || !w || !face_u || !face_v || !face_w) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc || (rc = vel_in(h, 0, {u, v, w, face_u, face_v, face_w}))) return rc;
    launch_face_vel(h);
    HIPCHK(hipGetLastError());
    return vel_out(h, 3, {face_u, face_v, face_w});
}

int fc_update_pressure(fc_pson *h, double *pPrime, double *p)
{
    if (check_handle(h) || !pPrime || !p) return fail(-1, "null argument");
    int rc = ensure_proj(h);
    if (rc || (rc = to_device(h, h->p, pPrime))) return rc;
    update_pressure(h->st, h->dims[0], h->pzero, h->p, h->pprev);
    HIPCHK(hipGetLastError());
    if ((rc = to_host(h, pPrime, h->p))) return rc;
    return to_host(h, p, h->pprev);
}

int fc_project_resident(fc_pson *h, double dt, double dt_fac)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    hipEvent_t ev[5] = {};
    if (h->timing)
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    auto mark = [&](int n) { if (h->timing) (void)hipEventRecord(ev[n], h->st); };
    mark(0);
    launch_correct_vel(h, h->p, dt);
    mark(1);
    launch_ghost_vel(h, h->p, dt_fac);
    mark(2);
    launch_face_vel(h);
    mark(3);
    update_pressure(h->st, h->dims[0], h->pzero, h->p, h->pprev);
    mark(4);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    if (h->timing) {
        for (int n = 0; n < 4; ++n) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[n], ev[n + 1]);
            h->proj_ms[n] = ms;
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    return 0;
}

int fc_face_vel_resident(fc_pson *h)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    launch_face_vel(h);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_pson_rhs_resident(fc_pson *h, double dtinv, double *sumd, double *resDiv)
{
    if (check_handle(h)) return -1;
    if (!h->have_masks) return fail(-1, "fc_pson_set_masks must be called before rhs_poisson");
    int rc = ensure_proj(h);
    if (rc) return rc;
    const LevelDims &d = h->dims[0];
    const bool mix = h->prm.mix_GC_form == 1;
    if (mix) {
        if ((rc = finalize_gcm(h))) return rc;
        if (h->gvel_dirty && (rc = prepare_gvel(h))) return rc;
    }
    double *du = vel_field(h, 3), *dv = vel_field(h, 4), *dw = vel_field(h, 5);
    // enforce_global_mass_consv corrects the resident face velocities in place
    if ((rc = mass_consv_device(h, du, dv, dw))) return rc;
    fill<double>(h->st, d.n, h->div, 0.0);
    rhs_poisson(h->st, d, du, dv, dw, h->igk[0], h->l1rhs, dtinv, h->div, h->partial, h->scal);
    if (mix) ghost_massflux(h->st, d, h->gvel, h->ngvel, du, dv, dw, h->l1rhs, dtinv, h->div, h->partial,
                            h->scal + 118);
    HIPCHK(hipMemcpyAsync(h->hscal, h->scal, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    if (mix) HIPCHK(hipMemcpyAsync(h->hscal + 118, h->scal + 118, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    if (sumd) *sumd = h->hscal[0] / dtinv + (mix ? h->hscal[118] : 0.0);
    if (resDiv) *resDiv = h->hscal[1];
    return 0;
}

int fc_pson_resident_copy(fc_pson *h, int which, double *dst)
{
    if (check_handle(h) || !dst) return fail(-1, "null argument");
    if (which < 0 || which > 7) return fail(-1, "which must be 0..7");
    const double *src = nullptr;
    if (which < 6) {
        int rc = ensure_proj(h);
        if (rc) return rc;
        src = vel_field(h, which);
    } else {
        src = which == 6 ? h->pprev : h->p;
    }
    const LevelDims &d = h->dims[0];
    const size_t w = sizeof(double) * ((size_t)d.nx + 2);
    HIPCHK(hipMemcpy2DAsync(dst, w, src, sizeof(double) * (size_t)d.px, w, (size_t)(d.ny + 2) * (d.nz + 2),
                            hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_project_timing_read(fc_pson *h, double *ms4)
{
    if (check_handle(h) || !ms4) return fail(-1, "null argument");
    for (int n = 0; n < 4; ++n) ms4[n] = h->proj_ms[n];
    return 0;
}

int fc_pson_download_p(fc_pson *h, double *p)
{
    if (check_handle(h) || !p) return fail(-1, "null argument");
    return to_host(h, p, h->pprev);
}

// ---------------------------------------------------------------------------
// Advection-diffusion line solves (CORE_SOLVE_AD.F90, CORE_SET_SOLVERS.F90).

namespace {

int ad_check(fc_pson *h)
{
    if (!h->ad_ready) return fail(-1, "advection-diffusion call before fc_set_solve_ad");
    return ensure_proj(h);
}

void launch_itsolv_ad(fc_pson *h, double *var, const double *r, int nsweep)
{
    const LevelDims &d = h->dims[0];
    for (int s = 0; s < nsweep; ++s)
        for (int dir = 0; dir < 3; ++dir)
            for (int c = 0; c < 2; ++c)
                ad_sweep(h->st, d, h->am, var, r, vel_field(h, 3), vel_field(h, 4), vel_field(h, 5), h->igk[0], dir,
                         c);
}

int read_ad_residual(fc_pson *h, const double *var, const double *r, double *resm, int *loc, double *res2)
{
    ad_residual(h->st, h->dims[0], h->am, var, r, vel_field(h, 3), vel_field(h, 4), vel_field(h, 5), h->igk[0],
                h->ad_partial, h->ad_partial + 3 * ROW_BLOCKS);
    HIPCHK(hipMemcpyAsync(h->hscal + 8, h->ad_partial + 3 * ROW_BLOCKS, 3 * sizeof(double), hipMemcpyDeviceToHost,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (resm) *resm = h->hscal[8];
    if (loc) {
        const long t = (long)h->hscal[9];
        const int ni = h->nx - 1, nj = h->ny - 1;
        if (t < 0) {
            loc[0] = loc[1] = loc[2] = 0;
        } else {
            loc[0] = 1 + (int)(t % ni);
            loc[1] = 1 + (int)((t / ni) % nj);
            loc[2] = 1 + (int)(t / ((long)ni * nj));
        }
    }
    if (res2) *res2 = h->hscal[10];
    return 0;
}

}  // namespace

int fc_set_solve_ad(fc_pson *h, double reinv, double dt, int crank_nicolson, double omega_ad)
{
    if (check_handle(h)) return -1;
    if (std::max(h->nx, std::max(h->ny, h->nz)) > 575)
        return fail(-2, "advection-diffusion lines longer than 576 rows are not supported by this build");
    int rc = ensure_proj(h);
    if (rc) return rc;
    // set_solve_ad (CORE_SET_SOLVERS.F90:591-676), nu = reinv on every face
    const int n3[3] = {h->nx, h->ny, h->nz};
    const HostMet &m = h->hm[0];
    const std::vector<double> *dcinv[3] = {&m.dxcinv, &m.dycinv, &m.dzcinv}, *dinv[3] = {&m.dxinv, &m.dyinv, &m.dzinv};
    size_t span[3], tot = 0;
    for (int a = 0; a < 3; ++a) { span[a] = (size_t)pad16(n3[a] + 2) + 16; tot += 2 * span[a]; }
    std::vector<double> hv(tot, 0.0);
    size_t off = 0, offs[6];
    for (int a = 0; a < 3; ++a) {
        offs[2 * a] = off;
        offs[2 * a + 1] = off + span[a];
        for (int i = 1; i <= n3[a] - 1; ++i) {
            double am = ((*dcinv[a])[i]) * (*dinv[a])[i];
            double ap = ((*dcinv[a])[i + 1]) * (*dinv[a])[i];
            hv[off + i] = -(0.50 * dt * reinv) * am;
            hv[off + span[a] + i] = -(0.50 * dt * reinv) * ap;
        }
        off += 2 * span[a];
    }
    if (!h->ad_coefs && (rc = dalloc(h, &h->ad_coefs, tot))) return rc;
    if (!h->ad_partial && (rc = dalloc(h, &h->ad_partial, 3 * (size_t)ROW_BLOCKS + 8))) return rc;
    if (!h->ad_rhs) {
        if ((rc = dalloc(h, &h->ad_rhs, 3 * (size_t)h->dims[0].n))) return rc;
        fill<double>(h->st, 3 * h->dims[0].n, h->ad_rhs, 0.0);
    }
    HIPCHK(hipMemcpyAsync(h->ad_coefs, hv.data(), sizeof(double) * tot, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    const double *c = h->ad_coefs;
    h->am = AdMet{h->pm, c + offs[0], c + offs[1], c + offs[2], c + offs[3], c + offs[4], c + offs[5],
                  crank_nicolson ? 0.5 * dt : 0.0, omega_ad};
    h->ad_reinv = reinv;
    h->ad_ready = true;
    return 0;
}

int fc_ad_upload(fc_pson *h, const double *nlu, const double *nlv, const double *nlw)
{
    if (check_handle(h)) return -1;
    int rc = ad_check(h);
    if (rc) return rc;
    const double *src[3] = {nlu, nlv, nlw};
    for (int c = 0; c < 3; ++c)
        if (src[c] && (rc = to_device(h, h->ad_rhs + (size_t)c * h->dims[0].n, src[c]))) return rc;
    return 0;
}

int fc_ad_download(fc_pson *h, double *nlu, double *nlv, double *nlw)
{
    if (check_handle(h)) return -1;
    int rc = ad_check(h);
    if (rc) return rc;
    double *dst[3] = {nlu, nlv, nlw};
    for (int c = 0; c < 3; ++c)
        if (dst[c] && (rc = to_host(h, dst[c], h->ad_rhs + (size_t)c * h->dims[0].n))) return rc;
    return 0;
}

int fc_itsolv_ad_resident(fc_pson *h, int comp, int nsweep)
{
    if (check_handle(h)) return -1;
    if (comp < 0 || comp > 2 || nsweep < 0) return fail(-1, "bad component or sweep count");
    int rc = ad_check(h);
    if (rc) return rc;
    hipEvent_t ev[2] = {};
    if (h->timing)
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    if (h->timing) HIPCHK(hipEventRecord(ev[0], h->st));
    launch_itsolv_ad(h, vel_field(h, comp), h->ad_rhs + (size_t)comp * h->dims[0].n, nsweep);
    if (h->timing) HIPCHK(hipEventRecord(ev[1], h->st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    if (h->timing) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        h->ad_ms = ms;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    return 0;
}

int fc_calc_residual_ad_resident(fc_pson *h, int comp, double *resm, int *loc, double *res2)
{
    if (check_handle(h)) return -1;
    if (comp < 0 || comp > 2) return fail(-1, "bad component");
    int rc = ad_check(h);
    if (rc) return rc;
    return read_ad_residual(h, vel_field(h, comp), h->ad_rhs + (size_t)comp * h->dims[0].n, resm, loc, res2);
}

int fc_itsolv_ad(fc_pson *h, double *var, const double *r, const double *face_u, const double *face_v,
                 const double *face_w, int nsweep)
{
    if (check_handle(h) || !var || !r || !face_u || !face_v || !face_w) return fail(-1, "null argument");
    int rc = ad_check(h);
    if (rc || (rc = vel_in(h, 0, {var})) || (rc = vel_in(h, 3, {face_u, face_v, face_w})) ||
        (rc = to_device(h, h->ad_rhs, r)))
        return rc;
    if ((rc = fc_itsolv_ad_resident(h, 0, nsweep))) return rc;
    return vel_out(h, 0, {var});
}

int fc_calc_residual_ad(fc_pson *h, const double *var, const double *r, const double *face_u, const double *face_v,
                        const double *face_w, double *resm, int *loc, double *res2)
{
    if (check_handle(h) || !var || !r || !face_u || !face_v || !face_w) return fail(-1, "null argument");
    int rc = ad_check(h);
    if (rc || (rc = vel_in(h, 0, {var})) || (rc = vel_in(h, 3, {face_u, face_v, face_w})) ||
        (rc = to_device(h, h->ad_rhs, r)))
        return rc;
    return read_ad_residual(h, vel_field(h, 0), h->ad_rhs, resm, loc, res2);
}

// ---------------------------------------------------------------------------
// solve_ad's loop (CORE_SOLVE_AD.F90:11-127) on the resident velocities:
// set_outer_velocity_bc, set_outer_ghost_velocity, GCM_ghostcell_velocity.

namespace {

VelBc make_velbc(fc_pson *h)
{
    VelBc b;
    for (int f = 0; f < 6; ++f) b.type[f] = h->prm.bc_vel[f];
    for (int t = 0; t < 18; ++t) { b.val[t] = h->vbc_val[t]; b.freq[t] = h->vbc_freq[t]; }
    b.y0 = h->vbc_y0;
    b.y1 = h->vbc_y1;
    b.pi = 4.0 * std::atan(1.0);
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) b.arr[3 * a + c] = vbc_ptr(h, a, c);
    return b;
}

int check_velbc_types(const fc_pson *h)
{
    for (int f = 0; f < 6; ++f) {
        const int t = h->prm.bc_vel[f];
        if (t < 1 || t > 7) return fail(-1, "bc_vel must be 1..7");
        if (t == 7) return fail(-2, "shear outer velocity faces are not implemented (the reference stops there, "
                                    "CORE_SET_OUTER_BC.F90:98-108)");
        if (t == 6 && f >= 2)
            return fail(-2, "user-specified velocity on y or z faces: the reference writes the x-face arrays there "
                            "(CORE_SET_OUTER_BC.F90:259-266), which this build does not reproduce");
    }
    return 0;
}

// enforce_global_mass_consv (CORE_SET_OUTER_BC.F90:515-735) on face
// velocities du, dv, dw on the device: no-op without a zero-gradient face
int mass_consv_device(fc_pson *h, double *du, double *dv, double *dw)
{
    const LevelDims &d = h->dims[0];
    const int zg = zg_mask(h);
    h->outflow_corr = 0.0;
    if (!zg) return 0;
    const bool mix = h->prm.mix_GC_form == 1;
    massflux(h->st, d, du, dv, dw, h->igk[0], h->l1rhs, h->partial, h->scal + 116);
    if (mix) ghost_massflux(h->st, d, h->gvel, h->ngvel, du, dv, dw, h->l1rhs, 1.0, nullptr, h->partial, h->scal + 117);
    else fill<double>(h->st, 1, h->scal + 117, 0.0);
    HIPCHK(hipMemcpyAsync(h->hscal + 116, h->scal + 116, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    const double mflux = h->hscal[116] + h->hscal[117];
    const double area = outflow_area_host(h, zg);
    if (!(area > 0.0)) return fail(-1, "enforce_global_mass_consv: the zero-gradient faces have no fluid area");
    h->outflow_area = area;
    h->outflow_corr = -mflux / area;
    outflow_velocity(h->st, d, zg, h->outflow_corr, du, dv, dw, h->velbc ? vbc_ptr(h, 0, 0) : nullptr,
                     h->velbc ? vbc_ptr(h, 1, 1) : nullptr, h->velbc ? vbc_ptr(h, 2, 2) : nullptr);
    return 0;
}

int gcm_velocity(fc_pson *h, int iopt, double restol_gcv, int *sweeps, double *resmax)
{
    if (!h->have_gcm_vel) return fail(-1, "GCM_ghostcell_velocity before fc_set_gcm_vel");
    int rc = finalize_gcm(h);
    if (rc) return rc;
    const bool mix = h->prm.mix_GC_form == 1;
    if (mix && !h->ad_ready) return fail(-1, "mix_GC_form = 1 velocity update before fc_set_solve_ad");
    if (mix)
        // the hybrid row reads the six neighbours of each ghost cell
        for (size_t n = 0; n < h->hg.iG.size(); ++n)
            if (h->hg.iG[n] < 1 || h->hg.iG[n] > h->nx - 1 || h->hg.jG[n] < 1 || h->hg.jG[n] > h->ny - 1 ||
                h->hg.kG[n] < 1 || h->hg.kG[n] > h->nz - 1)
                return fail(-1, "GCM_ghostcell_velocity with mix_GC_form = 1: ghost cell outside the interior");
    const GcmDev &g = h->g;
    const size_t nG = (size_t)g.nG, nD = (size_t)g.nD;
    GcmVel gv{h->gv_buf, h->gv_buf + 8 * nG, h->gv_buf + 8 * nG + 8 * nD, h->gv_buf + 8 * nG + 8 * nD + 3 * nG,
              h->gv_probe / (h->gv_probe - 1.0), 1.0 / (h->gv_probe - 1.0)};
    const long need = std::max<long>((g.nG + 255) / 256, (g.nD + 255) / 256);
    if (need > h->npartial) return fail(-1, "GCM list larger than the reduction scratch");
    int it = 0;
    double res = 1.0e10;
    const LevelDims &d = h->dims[0];
    while (it < h->prm.itermax_gc && res > restol_gcv) {
        gcm_velocity_step(h->st, g, gv, d, h->am, mix ? 1 : 0, iopt, vel_field(h, 0), vel_field(h, 1),
                          vel_field(h, 2), h->ad_rhs, h->ad_rhs ? h->ad_rhs + d.n : nullptr,
                          h->ad_rhs ? h->ad_rhs + 2 * d.n : nullptr, vel_field(h, 3), vel_field(h, 4),
                          vel_field(h, 5), h->gijk, h->off8, h->gv_vals, h->partial, h->scal + 120);
        HIPCHK(hipMemcpyAsync(h->hscal + 120, h->scal + 120, 2 * sizeof(double), hipMemcpyDeviceToHost, h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        res = std::max(h->hscal[120], h->hscal[121]);
        it++;
    }
    if (sweeps) *sweeps = it;
    if (resmax) *resmax = res;
    return (res > restol_gcv) ? 1 : 0;
}

}  // namespace

int fc_vel_set_outer(fc_pson *h, const double *vals, const double *freq, double y_origin, double y_out)
{
    if (check_handle(h)) return -1;
    for (int t = 0; t < 18; ++t) {
        h->vbc_val[t] = vals ? vals[t] : 0.0;
        h->vbc_freq[t] = freq ? freq[t] : 0.0;
    }
    h->vbc_y0 = y_origin;
    h->vbc_y1 = y_out;
    return 0;
}

int fc_outer_velocity_bc(fc_pson *h, double time)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc || (rc = check_velbc_types(h))) return rc;
    outer_velocity_bc(h->st, h->dims[0], make_velbc(h), vel_field(h, 0), vel_field(h, 1), vel_field(h, 2), h->d_ynode,
                      time);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_outer_ghost_velocity(fc_pson *h)
{
    if (check_handle(h)) return -1;
    int rc = ensure_proj(h);
    if (rc) return rc;
    outer_ghost_velocity(h->st, h->dims[0], make_velbc(h), vel_field(h, 0), vel_field(h, 1), vel_field(h, 2));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_set_gcm_vel(fc_pson *h, const double *coeffD, const double *coeffDeadD, const double *ubi,
                   const double *ubiDead, double probe_norm_d)
{
    if (check_handle(h)) return -1;
    const size_t nG = h->hg.iG.size(), nD = h->hg.iD.size();
    if ((nG && (!coeffD || !ubi)) || (nD && (!coeffDeadD || !ubiDead))) return fail(-1, "null argument");
    if (!(probe_norm_d > 1.0)) return fail(-1, "probeLengthNormalizedD must exceed 1");
    int rc = ensure_proj(h);
    if (rc) return rc;
    dfree(h, h->gv_buf);
    dfree(h, h->gv_vals);
    h->gv_buf = h->gv_vals = nullptr;
    const size_t nb = 8 * nG + 8 * nD + 3 * nG + 3 * nD;
    if ((rc = dalloc(h, &h->gv_buf, nb)) || (rc = dalloc(h, &h->gv_vals, 3 * std::max(nG, nD)))) return rc;
    std::vector<double> hb(nb, 0.0);
    if (nG) {
        memcpy(hb.data(), coeffD, sizeof(double) * 8 * nG);
        memcpy(hb.data() + 8 * nG + 8 * nD, ubi, sizeof(double) * 3 * nG);
    }
    if (nD) {
        memcpy(hb.data() + 8 * nG, coeffDeadD, sizeof(double) * 8 * nD);
        memcpy(hb.data() + 8 * nG + 8 * nD + 3 * nG, ubiDead, sizeof(double) * 3 * nD);
    }
    HIPCHK(hipMemcpyAsync(h->gv_buf, hb.data(), sizeof(double) * nb, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->gv_probe = probe_norm_d;
    h->have_gcm_vel = true;
    return 0;
}

int fc_ghostcell_velocity(fc_pson *h, int iopt, double restol_gcv, int *sweeps, double *resmax)
{
    if (check_handle(h)) return -1;
    if (iopt != 1 && iopt != 2) return fail(-1, "iOpt must be 1 or 2");
    int rc = ensure_proj(h);
    if (rc) return rc;
    return gcm_velocity(h, iopt, restol_gcv, sweeps, resmax);
}

int fc_solve_ad_resident(fc_pson *h, double time, int itermax_ad, double restol_ad, double restol_gcv, int *iters,
                         double *res, int *loc)
{
    if (check_handle(h)) return -1;
    int rc = ad_check(h);
    if (rc || (rc = check_velbc_types(h))) return rc;
    if (!h->have_gcm_vel && (h->hg.iG.size() || h->hg.iD.size()))
        return fail(-1, "solve_ad with ghost or dead cells before fc_set_gcm_vel");
    const LevelDims &d = h->dims[0];
    const VelBc vb = make_velbc(h);
    int iter = 0;
    double res2 = 1.0e10;
    int lc[3] = {0, 0, 0};
    while (iter < itermax_ad && res2 > restol_ad) {
        for (int c = 0; c < 3; ++c) launch_itsolv_ad(h, vel_field(h, c), h->ad_rhs + (size_t)c * d.n, 1);
        if (h->have_gcm_vel) {
            rc = gcm_velocity(h, 2, restol_gcv, nullptr, nullptr);
            if (rc > 0) return fail(-5, "GCM_ghostcell_velocity did not converge in itermax_gc sweeps");
            if (rc) return rc;
        }
        face_vel(h->st, d, h->pm, vel_field(h, 0), vel_field(h, 1), vel_field(h, 2), vbc_ptr(h, 0, 0),
                 vbc_ptr(h, 1, 1), vbc_ptr(h, 2, 2), vel_field(h, 3), vel_field(h, 4), vel_field(h, 5));
        if ((rc = mass_consv_device(h, vel_field(h, 3), vel_field(h, 4), vel_field(h, 5)))) return rc;
        outer_velocity_bc(h->st, d, vb, vel_field(h, 0), vel_field(h, 1), vel_field(h, 2), h->d_ynode, time);
        outer_ghost_velocity(h->st, d, vb, vel_field(h, 0), vel_field(h, 1), vel_field(h, 2));
        double r = -1.0;
        for (int c = 0; c < 3; ++c) {
            double rm = 0.0;
            int l3[3];
            if ((rc = read_ad_residual(h, vel_field(h, c), h->ad_rhs + (size_t)c * d.n, &rm, l3, nullptr))) return rc;
            rm = std::fabs(rm);
            if (rm > r) { r = rm; lc[0] = l3[0]; lc[1] = l3[1]; lc[2] = l3[2]; }
        }
        res2 = r;
        iter++;
    }
    HIPCHK(hipGetLastError());
    if (iters) *iters = iter;
    if (res) *res = res2;
    if (loc) { loc[0] = lc[0]; loc[1] = lc[1]; loc[2] = lc[2]; }
    return (res2 > restol_ad) ? 1 : 0;
}

int fc_solve_ad(fc_pson *h, double *u, double *v, double *w, const double *nlu, const double *nlv,
                const double *nlw, double *face_u, double *face_v, double *face_w, double time, int itermax_ad,
                double restol_ad, double restol_gcv, int *iters, double *res, int *loc)
{
    if (check_handle(h) || !u || !v || !w || !nlu || !nlv || !nlw || !face_u || !face_v || !face_w)
        return fail(-1, "null argument");
    int rc = ad_check(h);
    if (rc || (rc = vel_in(h, 0, {u, v, w, face_u, face_v, face_w})) || (rc = fc_ad_upload(h, nlu, nlv, nlw)))
        return rc;
    rc = fc_solve_ad_resident(h, time, itermax_ad, restol_ad, restol_gcv, iters, res, loc);
    if (rc < 0) return rc;
    int rc2 = vel_out(h, 0, {u, v, w, face_u, face_v, face_w});
    return rc2 ? rc2 : rc;
}

// ---------------------------------------------------------------------------
// rhs_advec_diff (CORE_SET_SOLVERS.F90:10-61) on the resident velocities.

namespace {

// the second-order upwind weights of rhs_advec per axis (CORE_SET_SOLVERS.F90:
// 105-121, 264-276, 355-369), from level 1's inverse spacings (the grid's,
// CORE_GRID.F90:120-139); each rounded on its own, as in the reference
int ensure_upw(fc_pson *h)
{
#pragma clang fp contract(off)
    if (h->upw) return 0;
    const int n3[3] = {h->nx, h->ny, h->nz};
    const HostMet &m = h->hm[0];
    const std::vector<double> *dinv[3] = {&m.dxinv, &m.dyinv, &m.dzinv};
    size_t tot = 0;
    for (int a = 0; a < 3; ++a) tot += (size_t)n3[a] + 2;
    std::vector<UpwW> hv(tot, UpwW{});
    size_t off = 0;
    for (int a = 0; a < 3; ++a) {
        const std::vector<double> &d = *dinv[a];
        const int N = n3[a];
        for (int n = 1; n <= N - 1; ++n) {
            const int nm = std::max(n - 2, 0), np = std::min(n + 2, N);
            UpwW &w = hv[off + n];
            w.em1 = (d[n - 1] + 2.0 * d[n]) / (d[n - 1] + d[n]);
            w.em2 = d[n] / (d[n - 1] + d[n]);
            w.ep1 = (d[np] + 2.0 * d[n + 1]) / (d[np] + d[n + 1]);
            w.ep2 = d[n + 1] / (d[np] + d[n + 1]);
            w.wm1 = (d[nm] + 2.0 * d[n - 1]) / (d[nm] + d[n - 1]);
            w.wm2 = d[n - 1] / (d[nm] + d[n - 1]);
            w.wp1 = (d[n + 1] + 2.0 * d[n]) / (d[n + 1] + d[n]);
            w.wp2 = d[n] / (d[n + 1] + d[n]);
        }
        off += (size_t)N + 2;
    }
    int rc = dalloc(h, &h->upw, tot);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(h->upw, hv.data(), sizeof(UpwW) * tot, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

RhsAdMet rhs_met(fc_pson *h, double alfa, double dt)
{
    const UpwW *w = h->upw;
    return RhsAdMet{h->pm, h->l1met[0], h->l1met[1], h->l1met[2], w, w + h->nx + 2, w + h->nx + 2 + h->ny + 2,
                    h->have_visc ? h->visc_tot : nullptr, h->ad_reinv, alfa, dt};
}

}  // namespace

int fc_set_visc_tot(fc_pson *h, const double *viscTot)
{
    if (check_handle(h)) return -1;
    int rc = ad_check(h);
    if (rc) return rc;
    if (!viscTot) {
        h->have_visc = false;
        return 0;
    }
    if (!h->visc_tot && (rc = dalloc(h, &h->visc_tot, (size_t)h->dims[0].n))) return rc;
    if ((rc = to_device(h, h->visc_tot, viscTot))) return rc;
    h->have_visc = true;
    return 0;
}

int fc_rhs_advec_diff_resident(fc_pson *h, double alfa_upwind, double dt)
{
    if (check_handle(h)) return -1;
    int rc = ad_check(h);
    if (rc || (rc = ensure_upw(h))) return rc;
    const size_t n = (size_t)h->dims[0].n;
    hipEvent_t ev[2] = {};
    if (h->timing) {
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(ev[0], h->st));
    }
    rhs_advec_diff(h->st, h->dims[0], rhs_met(h, alfa_upwind, dt), vel_field(h, 0), vel_field(h, 1), vel_field(h, 2),
                   vel_field(h, 3), vel_field(h, 4), vel_field(h, 5), h->ad_rhs, h->ad_rhs + n, h->ad_rhs + 2 * n);
    if (h->timing) (void)hipEventRecord(ev[1], h->st);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(h->st);
    if (h->timing) {
        float ms = 0.0f;
        if (le == hipSuccess && se == hipSuccess) (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
        h->ad_ms = ms;
        for (auto &e : ev) (void)hipEventDestroy(e);
    }
    HIPCHK(le);
    HIPCHK(se);
    return 0;
}

int fc_rhs_advec_diff(fc_pson *h, const double *uOld, const double *vOld, const double *wOld, const double *face_u,
                      const double *face_v, const double *face_w, double alfa_upwind, double dt, double *nlu,
                      double *nlv, double *nlw)
{
    if (check_handle(h) || !uOld || !vOld || !wOld || !face_u || !face_v || !face_w || !nlu || !nlv || !nlw)
        return fail(-1, "null argument");
    int rc = ad_check(h);
    if (rc || (rc = vel_in(h, 0, {uOld, vOld, wOld, face_u, face_v, face_w})) ||
        (rc = fc_ad_upload(h, nlu, nlv, nlw)) || (rc = fc_rhs_advec_diff_resident(h, alfa_upwind, dt)))
        return rc;
    return fc_ad_download(h, nlu, nlv, nlw);
}

int fc_ad_timing_read(fc_pson *h, double *ms)
{
    if (check_handle(h) || !ms) return fail(-1, "null argument");
    *ms = h->ad_ms;
    return 0;
}

int fc_pson_timing(fc_pson *h, int enable)
{
    if (check_handle(h)) return -1;
    h->timing = enable != 0;
    h->sweep_ms = 0.0;
    h->sweep_launches = 0;
    h->ev_used = 0;
    return 0;
}

int fc_pson_timing_read(fc_pson *h, double *sweep_ms, long *launches)
{
    if (check_handle(h)) return -1;
    timing_flush(h);
    if (sweep_ms) *sweep_ms = h->sweep_ms;
    if (launches) *launches = h->sweep_launches;
    return 0;
}

// ---- single level operations ---------------------------------------------
// prec: the hierarchy (FC_PREC_FP64 / FC_PREC_MIXED) and FC_LVL_WALLS
static int lvl_check(fc_pson *h, int lev, int prec)
{
    if (check_handle(h)) return -1;
    if (lev < 1 || lev > h->nlev) return fail(-1, "level out of range");
    const int pr = prec & ~FC_LVL_WALLS;
    if (pr != FC_PREC_FP64 && pr != FC_PREC_MIXED) return fail(-1, "bad precision");
    if (pr == FC_PREC_MIXED) return ensure_mixed(h);
    return 0;
}

int fc_lvl_write(fc_pson *h, int lev, int which, const void *host)
{
    int rc = lvl_check(h, lev, 0);
    if (rc) return rc;
    int l = lev - 1;
    switch (which) {
    case 0: rc = to_device(h, h->L64[l].phi, (const double *)host, l); break;
    case 1: rc = to_device(h, h->L64[l].rhs, (const double *)host, l); break;
    case 2: rc = to_device(h, h->L64[l].res, (const double *)host, l); break;
    case 3: rc = to_device(h, h->ibk[l], (const int8_t *)host, l); break;
    case 4: rc = to_device(h, h->igk[l], (const int8_t *)host, l); break;
    case 10: case 11: case 12:
        if ((rc = ensure_mixed(h))) return rc;
        rc = to_device(h, which == 10 ? h->L32[l].phi : which == 11 ? h->L32[l].rhs : h->L32[l].res, (const float *)host, l);
        break;
    default: return fail(-1, "bad array selector");
    }
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_read(fc_pson *h, int lev, int which, void *host)
{
    int rc = lvl_check(h, lev, 0);
    if (rc) return rc;
    int l = lev - 1;
    switch (which) {
    case 0: rc = to_host(h, (double *)host, h->L64[l].phi, l); break;
    case 1: rc = to_host(h, (double *)host, h->L64[l].rhs, l); break;
    case 2: rc = to_host(h, (double *)host, h->L64[l].res, l); break;
    case 3: rc = to_host(h, (int8_t *)host, h->ibk[l], l); break;
    case 4: rc = to_host(h, (int8_t *)host, h->igk[l], l); break;
    case 5: rc = to_host(h, (int8_t *)host, h->wmask[l], l); break;
    case 6:
        if (l == 0) return fail(-1, "level 1 has no prolongation flags");
        rc = to_host(h, (int8_t *)host, h->wpflag[l], l);
        break;
    case 10: case 11: case 12:
        if ((rc = ensure_mixed(h))) return rc;
        rc = to_host(h, (float *)host, which == 10 ? h->L32[l].phi : which == 11 ? h->L32[l].rhs : h->L32[l].res, l);
        break;
    default: return fail(-1, "bad array selector");
    }
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_smooth(fc_pson *h, int lev, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    const bool walls = (prec & FC_LVL_WALLS) != 0;
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED) smooth<float>(h, h->L32, h->m32, lev - 1, false, true, walls);
    else smooth<double>(h, h->L64, h->m64, lev - 1, false, true, walls);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_sweep(fc_pson *h, int lev, int dir, int color, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    if (dir < 0 || dir > 2 || color < 0 || color > 1) return fail(-1, "bad direction or colour");
    int l = lev - 1;
    const LevelDims &d = h->dims[l];
    BcInfo bc = bcinfo(h, l, false);
    const int8_t *mask = (prec & FC_LVL_WALLS) ? h->wmask[l] : h->igk[l];
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED) {
        bc_fill<float>(h->st, d, h->L32[l].phi, bc);
        sweep<float>(h->st, d, h->m32[l].m, h->L32[l].phi, h->L32[l].rhs, mask, dir, color, h->prm.omega_pson, bc);
        if (bc.per) bc_fill<float>(h->st, d, h->L32[l].phi, bc);
    } else {
        bc_fill<double>(h->st, d, h->L64[l].phi, bc);
        sweep<double>(h->st, d, h->m64[l].m, h->L64[l].phi, h->L64[l].rhs, mask, dir, color, h->prm.omega_pson, bc);
        if (bc.per) bc_fill<double>(h->st, d, h->L64[l].phi, bc);
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_coarse_solve(fc_pson *h, int prec)
{
    int rc = lvl_check(h, h ? h->nlev : 1, prec & ~FC_LVL_WALLS);
    if (rc) return rc;
    if ((rc = ensure_coarse_direct(h))) return rc;
    if (h->cd_state != 1)
        return fail(-2, "coarse direct solve not in use (coarse_solve = 1 and vc_walls >= 1 with at most "
                        "FC_COARSE_DIRECT_MAX free coarsest cells)");
    const int l = h->nlev - 1;
    BcInfo bc = bcinfo(h, l, true);
    if ((prec & ~FC_LVL_WALLS) == FC_PREC_MIXED) {
        coarse_direct<float>(h->st, h->cd_n, h->cd_M, h->cd_q, h->L32[l].rhs, h->L32[l].phi);
        bc_fill<float>(h->st, h->dims[l], h->L32[l].phi, bc);
    } else {
        coarse_direct<double>(h->st, h->cd_n, h->cd_M, h->cd_q, h->L64[l].rhs, h->L64[l].phi);
        bc_fill<double>(h->st, h->dims[l], h->L64[l].phi, bc);
    }
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_bc(fc_pson *h, int lev, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    int l = lev - 1;
    BcInfo bc = bcinfo(h, l, false);
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED) bc_fill<float>(h->st, h->dims[l], h->L32[l].phi, bc);
    else bc_fill<double>(h->st, h->dims[l], h->L64[l].phi, bc);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_residual(fc_pson *h, int lev, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    int l = lev - 1;
    const int8_t *mask = (prec & FC_LVL_WALLS) ? h->wmask[l] : h->igk[l];
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED)
        residual<float>(h->st, h->dims[l], h->m32[l].m, h->L32[l].phi, h->L32[l].rhs, h->L32[l].res, mask);
    else
        residual<double>(h->st, h->dims[l], h->m64[l].m, h->L64[l].phi, h->L64[l].rhs, h->L64[l].res, mask);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_restrict(fc_pson *h, int lev, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    if (lev >= h->nlev) return fail(-1, "no coarser level");
    int l = lev - 1;
    const int8_t *rf = (prec & FC_LVL_WALLS) ? nullptr : h->rflag[l + 1];
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED)
        restrict_<float>(h->st, h->dims[l], h->dims[l + 1], h->m32[l].m, h->m32[l + 1].m, h->L32[l].res,
                         h->L32[l + 1].rhs, h->ibk[l], rf);
    else
        restrict_<double>(h->st, h->dims[l], h->dims[l + 1], h->m64[l].m, h->m64[l + 1].m, h->L64[l].res,
                          h->L64[l + 1].rhs, h->ibk[l], rf);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

int fc_lvl_prolong(fc_pson *h, int lev, int prec)
{
    int rc = lvl_check(h, lev, prec);
    if (rc) return rc;
    if (lev < 2) return fail(-1, "no finer level");
    int l = lev - 1;
    const VcMasks vm = vc_masks(h, (prec & FC_LVL_WALLS) != 0);
    prec &= ~FC_LVL_WALLS;
    if (prec == FC_PREC_MIXED)
        prolong<float>(h->st, h->dims[l - 1], h->dims[l], h->m32[l - 1].m, h->m32[l].m, h->L32[l].phi,
                       h->L32[l - 1].phi, (*vm.ibk)[l - 1], (*vm.ibk)[l], (*vm.pflag)[l], vm.walls);
    else
        prolong<double>(h->st, h->dims[l - 1], h->dims[l], h->m64[l - 1].m, h->m64[l].m, h->L64[l].phi,
                        h->L64[l - 1].phi, (*vm.ibk)[l - 1], (*vm.ibk)[l], (*vm.pflag)[l], vm.walls);
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

}  // extern "C"

// ===========================================================================
// Internal interface for the request server (fc_handle.h)
namespace fc {

int set_error(int code, const std::string &msg) { return fail(code, msg); }

int Arena::init(size_t bytes)
{
    bytes = (bytes + 255) / 256 * 256;
    if (hipMalloc((void **)&base_, bytes) != hipSuccess) {
        base_ = nullptr;
        return fail(-4, "hipMalloc of the device memory pool failed");
    }
    cap_ = bytes;
    used_ = 0;
    free_.clear();
    live_.clear();
    free_[0] = bytes;
    return 0;
}

void Arena::release()
{
    if (base_) hipFree(base_);
    base_ = nullptr;
    cap_ = used_ = 0;
    free_.clear();
    live_.clear();
}

void *Arena::alloc(size_t bytes)
{
    std::lock_guard<std::mutex> lk(mu_);
    bytes = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        size_t off = it->first, sz = it->second;
        free_.erase(it);
        if (sz > bytes) free_[off + bytes] = sz - bytes;
        live_[off] = bytes;
        used_ += bytes;
        return base_ + off;
    }
    return nullptr;
}

void Arena::free(void *p)
{
    std::lock_guard<std::mutex> lk(mu_);
    if (!p || !owns(p)) return;
    size_t off = (size_t)((char *)p - base_);
    auto lv = live_.find(off);
    if (lv == live_.end()) return;
    size_t sz = lv->second;
    live_.erase(lv);
    used_ -= sz;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + sz == nx->first) {      // merge with the next free piece
        sz += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {                             // and with the previous one
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += sz;
            return;
        }
    }
    free_[off] = sz;
}

void handle_set_stream(fc_pson *h, hipStream_t st)
{
    if (h->st && h->own_stream) hipStreamDestroy(h->st);
    h->st = st;
    h->own_stream = false;
}

int handle_create(const fc_grid *gr, const fc_pson_params *p, Arena *pool, hipStream_t stream, fc_pson **out)
{
    return create_impl(gr, p, pool, stream, out);
}

size_t handle_footprint(const fc_grid *gr, const fc_pson_params *p, int nGhost, int nDead)
{
    auto a = [](size_t b) { return (std::max<size_t>(b, 1) + 255) / 256 * 256; };
    int mgx = auto_levels(gr->nx, p->mgrids_x), mgy = auto_levels(gr->ny, p->mgrids_y), mgz = auto_levels(gr->nz, p->mgrids_z);
    int ml = std::max(mgx, std::max(mgy, mgz));
    const bool mixed = p->precision == FC_PREC_MIXED;
    int I = gr->nx, J = gr->ny, K = gr->nz;
    size_t tot = 0, n0 = 0;
    for (int l = 0; l < ml; ++l) {
        if (l > 0) {
            I = (l + 1 <= mgx) ? coarsen(I) : I;
            J = (l + 1 <= mgy) ? coarsen(J) : J;
            K = (l + 1 <= mgz) ? coarsen(K) : K;
        }
        size_t n = (size_t)pad16(I + 2) * (J + 2) * (K + 2);
        if (l == 0) n0 = n;
        auto span = [](int n) { return (size_t)((n + 3) / 4 * 4 + 4); };   // as upload_metrics
        size_t met = 7 * (span(I + 2) + span(J + 2) + span(K + 2));
        tot += a(8 * met) + a(4 * met) + 3 * a(n) + (l > 0 ? 3 * a(n) : 0) + a(8 * n) + (l > 0 ? 2 * a(8 * n) : 0);
        if (mixed) tot += 3 * a(4 * n);
    }
    size_t nx2 = gr->nx + 2, ny2 = gr->ny + 2, nz2 = gr->nz + 2;
    tot += 4 * a(8 * n0);                                                     // p, div, pprev, p_temp
    tot += 2 * a(8 * ny2 * nz2) + 2 * a(8 * nx2 * nz2) + 2 * a(8 * nx2 * ny2);  // pgrad planes
    tot += a(8 * 9 * (nx2 + ny2 + nz2));
    long np = std::max<long>((long)ny2 * nz2, 8192);
    np = std::max<long>(np, (long)(n0 / 256 + 1024));
    np = std::max<long>(np, (long)ROW_COMPONENTS * ROW_BLOCKS / 2);
    tot += a(16 * (size_t)np) + a(8 * 128);
    tot += a(360 * (size_t)std::max(nGhost, 1) + 400 * (size_t)std::max(nDead, 1) + 4096);   // GCM lists, records
    tot += a(4 * (size_t)(nGhost + nDead + 2));                                               // level offsets
    const int m = p->aa_depth > 0 ? std::min(p->aa_depth, 8) : 8;
    tot += 2 * a(8 * n0) + 2 * a(4 * n0) + 2 * (size_t)m * a(4 * n0) + a(64);
    return tot;
}

int handle_describe(fc_pson *h, ReqDesc *desc, ReqShape *shape)
{
    if (!h || !desc || !shape) return fail(-1, "null argument");
    const fc_pson_params &P = h->prm;
    for (int f = 0; f < 6; ++f)
        if (P.bc_vel[f] == FC_BC_TYPE_ZERO_GRADIENT || P.bc_vel[f] == FC_BC_TYPE_PERIODIC)
            return fail(-2, "zero-gradient and periodic outer faces are not supported by this build");
    if (P.it_solver_type != FC_IT_SOLVER_TYPE_MG) return fail(-2, "only it_solver_type = 3 (MG) runs on the GPU");
    if (P.outer_rule != FC_OUTER_RELRES || P.restol_pson > 0.0 || P.iterInter != 0 || P.fp32_cycles > 1 ||
        P.mix_GC_form == 1)
        return fail(-2, "batched solves need outer_rule RELRES, restol_pson <= 0, iterInter = 0, fp32_cycles <= 1 "
                        "and mix_GC_form = 0 (no host decision inside an outer iteration)");
    if (P.coarse_solve == 1) return fail(-2, "the coarsest-level direct solve is not part of batched solves");
    const bool mixed = P.precision == FC_PREC_MIXED;
    // batched FP64 requests run their V-cycles on p itself (the reference's
    // form); the wall correction cycles are the mixed requests'
    if (!mixed) h->prm.vc_walls = 0;
    int rc = finalize_gcm(h);
    if (rc) return rc;
    const bool gcm = P.itermax_gc > 0;
    if (gcm && !h->gsched_ok) return fail(-2, "batched solves need an acyclic GCM dependency graph");
    if (mixed && (rc = ensure_mixed(h))) return rc;
    const int m = P.aa_depth > 0 ? std::min(P.aa_depth, 8) : 8;
    if ((rc = ensure_aa(h, m)) || (rc = ensure_aa_fields(h))) return rc;
    if (mixed && !mixed_fold(h)) return fail(-2, "batched mixed-precision solves need the exact GCM pass");
    std::vector<int> gl = gcm ? h->glevel_off : std::vector<int>{0};
    if (!h->d_glevel) {
        if ((rc = dalloc(h, &h->d_glevel, gl.size()))) return rc;
        HIPCHK(hipMemcpyAsync(h->d_glevel, gl.data(), sizeof(int) * gl.size(), hipMemcpyHostToDevice, h->st));
    }
    if (h->nlev > MAX_LEVELS) return fail(-1, "more levels than max_grid_level");
    ReqDesc &D = *desc;
    memset(&D, 0, sizeof(D));
    D.nlev = h->nlev;
    D.mixed = mixed ? 1 : 0;
    const bool walls = mixed && P.vc_walls >= 1;
    D.walls = walls ? 1 : 0;
    D.homo = P.is_PBC_homogeneous == 1;
    D.nrb = ROW_BLOCKS;
    D.q111 = pin_index(h);
    shape->nlev = h->nlev;
    for (int l = 0; l < h->nlev; ++l) {
        ReqLevel &L = D.L[l];
        L.d = h->dims[l];
        L.m32 = h->m32[l].m;
        L.m64 = h->m64[l].m;
        L.phi32 = mixed ? h->L32[l].phi : nullptr;
        L.rhs32 = mixed ? h->L32[l].rhs : nullptr;
        L.phi64 = h->L64[l].phi;
        L.rhs64 = h->L64[l].rhs;
        L.igk = h->igk[l];
        L.ibk = h->ibk[l];
        L.rflag = h->rflag[l];
        L.pflag = h->pflag[l];
        // the masks of the request's V-cycles (walls: mixed correction cycles)
        const VcMasks vm = vc_masks(h, walls);
        L.vrow = (*vm.row)[l];
        L.vibk = (*vm.ibk)[l];
        L.vrflag = vm.rflag ? (*vm.rflag)[l] : nullptr;
        L.vpflag = (*vm.pflag)[l];
        for (int dir = 0; dir < 3; ++dir) {
            L.sw[dir] = sweep_shape(h->dims[l], dir, 0, P.omega_pson);
            if (L.sw[dir].R > 512) return fail(-1, "batched solves take at most 513 nodes per axis");
            shape->sw[l][dir] = L.sw[dir];
        }
        shape->d[l] = h->dims[l];
    }
    D.bc[0] = bcinfo(h, 0, false);
    D.bc[1] = bcinfo(h, 0, true);
    D.g = h->g;
    D.gnode = h->gnode;
    D.glevel = h->d_glevel;
    D.gdelta = (mixed && gcm) ? h->gdelta : nullptr;
    D.ntgt = gl.back();
    D.off8 = h->off8;
    D.p = h->p;
    D.div = h->div;
    D.pprev = h->pprev;
    D.aa_x = h->aa_x;
    D.aa_g = h->aa_g;
    D.aa_fprev = h->aa_fprev;
    D.aa_c = h->aa_c;
    D.cols = h->aa_cols;
    D.ring = h->aa_ring;
    D.partial = h->partial;
    D.scal = h->scal;
    shape->glevel = gl;
    shape->nrb = D.nrb;
    shape->cycle = cycle_shape(h);
    shape->mixed = mixed;
    HIPCHK(hipStreamSynchronize(h->st));
    return 0;
}

}  // namespace fc
