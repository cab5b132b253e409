// fc_serve.hip -- request server over the gfx950 pressure-Poisson path
// (include/fc_poisson.h, fc_serve_*): admission by device memory, resident
// hierarchies in one device pool, and continuous batching -- every step runs
// one accelerated outer iteration of solve_poisson (CORE_POISSON.F90:83-200
// with Anderson mixing) for all admitted requests, each phase of it as one
// batched launch (fc_kernels.hip, kb_*).  Requests join between steps and
// leave when they meet their own stop test.
//
// One host thread drives a server (submit / step / poll).  Admission runs on
// worker threads: a worker takes the oldest waiting request once the pool
// can hold its footprint, builds its handle (metrics, level masks, GCM
// schedule, uploads) on the worker's own stream, and hands it over ready;
// the next step gives it a slot.  The batched phases run on the execution
// stream, which is idle at every step boundary (each step ends with the
// read-back of the stop-test sums).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fc_poisson.h"
#include "fc_handle.h"
#include "fc_kernels.h"

using namespace fc;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id = 0;
    fc_serve_request rq{};
    Clock::time_point t_submit;
    size_t need = 0;               // footprint bound (handle_footprint)
    fc_pson *h = nullptr;
    int slot = -1;
    ReqDesc desc{};                // host copy of the device view
    ReqShape shape;
    bool joining = false;
    int iter = 0;
    int rc = 0;
    std::string err;
    fc_pson_stats st{};
};

struct Done {
    long long id;
    int rc;
    fc_pson_stats st;
};

// one batched launch: entries [off, off+ne) of the step's entry array
struct Launch {
    std::function<void(const BEnt *, int, int)> fn;
    size_t off;
    int ne, nb;
};

#define SCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return set_error(-3, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
    } while (0)

constexpr int ENT_CAP = 1 << 17;

}  // namespace

struct fc_serve {
    fc_serve_config cfg{};
    hipStream_t xs = nullptr, as = nullptr;   // execution; downloads and releases
    Arena pool;
    ReqDesc *d_desc = nullptr, *h_desc = nullptr;
    BEnt *d_ent = nullptr, *h_ent = nullptr;
    double *d_out = nullptr, *h_out = nullptr;
    // shared with the admission workers (under mu)
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request *> queue;     // submitted, not yet taken by a worker
    std::deque<Request *> ready;     // handle built, waiting for a slot
    std::deque<Request *> failed;    // admission errors
    size_t committed = 0;            // footprints of requests holding (or building) a handle
    long long releases = 0;          // handles released so far
    int preparing = 0;
    bool stop = false;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> wst;
    // driving thread only
    std::vector<Request *> active;
    std::vector<int> free_slots;
    std::deque<Done> done;
    long long next_id = 1, steps = 0;
    int last_launches = 0;
    // the step being built
    std::vector<BEnt> ent;
    std::vector<Launch> launches;
};

namespace {

// ---------------------------------------------------------------------------
// step program: entries of every phase of one outer iteration
struct Builder {
    fc_serve *s;
    // add one phase over the requests rs; blocks(r) gives r's block grid
    void phase(const std::vector<Request *> &rs, const std::function<bool(Request *)> &take,
               const std::function<dim3(Request *)> &blocks, std::function<void(const BEnt *, int, int)> fn,
               int align = 1)
    {
        const size_t off = s->ent.size();
        int nb = 0;
        for (Request *r : rs) {
            if (!take(r)) continue;
            dim3 g = blocks(r);
            const int n = (int)(g.x * g.y * g.z);
            if (n <= 0) continue;
            nb = (nb + align - 1) / align * align;
            s->ent.push_back(BEnt{nb, (int)g.x, (int)g.y, (int)g.z, r->slot});
            nb += n;
        }
        const int ne = (int)(s->ent.size() - off);
        if (ne > 0) s->launches.push_back(Launch{std::move(fn), off, ne, nb});
    }
};

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

long face_cells(const LevelDims &d)
{
    return (long)(d.ny - 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.nz - 1) + (long)(d.nx + 1) * (d.ny - 1);
}

// smoothing of level l (itsolv_mg: BC fill, then x, y, z lines in two colours)
// fill = false: the level's outer ghosts were zeroed by the fused restriction
// into it (its interior is zero), which is what the boundary fill would set
void smooth_phases(Builder &B, fc_serve *s, const std::vector<Request *> &grp, int prec, int l,
                   const std::function<bool(Request *)> &take, bool fill = true)
{
    const ReqDesc *D = s->d_desc;
    hipStream_t xs = s->xs;
    const int bcsel = (l == 0 && prec == 0) ? 0 : 1;   // FP64 level 1 carries the given outer BC
    if (fill)
        B.phase(grp, take, [l](Request *r) { return dim3(cdiv(face_cells(r->shape.d[l]), 256)); },
                [xs, D, prec, l, bcsel](const BEnt *E, int ne, int nb) {
                    b_bc_fill(xs, E, ne, nb, D, prec, l, 0, bcsel);
                });
    size_t lds = 0;
    for (Request *r : grp)
        if (take(r))
            for (int dir = 0; dir < 3; ++dir)
                lds = std::max(lds, sweep_lds_bytes(r->shape.sw[l][dir], prec ? sizeof(float) : sizeof(double)));
    for (int dir = 0; dir < 3; ++dir)
        for (int c = 0; c < 2; ++c)
            B.phase(grp, take, [l, dir](Request *r) { return dim3(r->shape.sw[l][dir].nblocks); },
                    [xs, D, prec, l, dir, c, bcsel, lds](const BEnt *E, int ne, int nb) {
                        b_sweep(xs, E, ne, nb, lds, D, prec, l, dir, c, bcsel);
                    },
                    8);
}

// one V-cycle of mg_solve (CORE_MG.F90:1618-1766) for a group of one precision;
// requests with fewer levels skip the phases of levels they do not have
void vcycle_phases(Builder &B, fc_serve *s, const std::vector<Request *> &grp, int prec)
{
    if (grp.empty()) return;
    const ReqDesc *D = s->d_desc;
    hipStream_t xs = s->xs;
    int lmax = 0;
    for (Request *r : grp) lmax = std::max(lmax, r->shape.nlev);
    for (int l = 0; l < lmax; ++l) {
        // the FP32 correction of a mixed request starts at zero with its
        // outer ghosts (kb_resid64_to32), like every level below level 1
        smooth_phases(B, s, grp, prec, l, [l](Request *r) { return r->shape.nlev > l; }, l == 0 && prec == 0);
        B.phase(grp, [l](Request *r) { return r->shape.nlev > l + 1; },
                [l](Request *r) {
                    const LevelDims &c = r->shape.d[l + 1];
                    return dim3(cdiv(c.nx - 1, 64), cdiv(c.ny - 1, 4), c.nz - 1);
                },
                [xs, D, prec, l](const BEnt *E, int ne, int nb) { b_resrestrict(xs, E, ne, nb, D, prec, l); });
    }
    for (int l = lmax - 1; l >= 1; --l) {
        smooth_phases(B, s, grp, prec, l, [l](Request *r) { return l < r->shape.nlev - 1; });
        B.phase(grp, [l](Request *r) { return l < r->shape.nlev; },
                [l](Request *r) { return prolong_grid(r->shape.d[l - 1], r->shape.d[l]); },
                [xs, D, prec, l](const BEnt *E, int ne, int nb) { b_prolong(xs, E, ne, nb, D, prec, l); });
    }
    smooth_phases(B, s, grp, prec, 0, [](Request *r) { return r->shape.cycle == 2; });
}

// give back a request's handle and slot (driving thread)
void release(fc_serve *s, Request *r)
{
    if (r->h) {
        fc_pson_destroy(r->h);
        r->h = nullptr;
        std::lock_guard<std::mutex> lk(s->mu);
        s->committed -= r->need;
        s->releases++;
        s->cv.notify_all();
    }
    if (r->slot >= 0) s->free_slots.push_back(r->slot);
    r->slot = -1;
}

void finish(fc_serve *s, Request *r, int rc)
{
    r->rc = rc;
    r->st.time_ms = std::chrono::duration<double, std::milli>(Clock::now() - r->t_submit).count();
    s->done.push_back(Done{r->id, rc, r->st});
    release(s, r);
    delete r;
}

// build and describe the handle of a request on stream st (worker thread)
int prepare(fc_serve *s, Request *r, hipStream_t st)
{
    const fc_serve_request &q = r->rq;
    fc_pson *h = nullptr;
    int rc = handle_create(&q.grid, &q.params, &s->pool, st, &h);
    if (rc) return rc;
    r->h = h;
    if ((rc = fc_pson_set_masks(h, q.iblank, q.ghost, q.dead))) return rc;
    if ((rc = fc_pson_set_gcm(h, q.nGhost, q.iGhost, q.jGhost, q.kGhost, q.iCellIndex, q.jCellIndex, q.kCellIndex,
                              q.coeffGCMN, q.alphaGhost, q.dpdnBodyIntercept, q.nDead, q.iDead, q.jDead, q.kDead,
                              q.iDeadCellIndex, q.jDeadCellIndex, q.kDeadCellIndex, q.coeffGCMDeadN,
                              q.probeLengthDead, q.dpdnInterceptDead, q.xyzDCIP)))
        return rc;
    if ((rc = fc_pson_set_bc(h, q.pgrad[0], q.pgrad[1], q.pgrad[2], q.pgrad[3], q.pgrad[4], q.pgrad[5], 1.0)))
        return rc;
    if ((rc = fc_pson_upload(h, nullptr, q.div, nullptr))) return rc;
    if ((rc = handle_describe(h, &r->desc, &r->shape))) return rc;
    SCHK(hipStreamSynchronize(st));
    handle_set_stream(h, s->as);
    return 0;
}

// admission worker: the oldest waiting request, once the pool can hold it
void worker(fc_serve *s, int w)
{
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv.wait(lk, [s] {
            return s->stop ||
                   (!s->queue.empty() && (s->committed + s->queue.front()->need <= s->pool.capacity() ||
                                          s->queue.front()->need > s->pool.capacity()));
        });
        if (s->stop) return;
        Request *r = s->queue.front();
        s->queue.pop_front();
        if (r->need > s->pool.capacity()) {
            r->rc = -6;
            r->err = "request needs more device memory than the pool holds";
            s->failed.push_back(r);
            continue;
        }
        s->committed += r->need;
        s->preparing++;
        const long long rel0 = s->releases;
        lk.unlock();
        int rc = prepare(s, r, s->wst[w]);
        std::string err = rc ? std::string(fc_pson_last_error()) : std::string();
        if (rc && r->h) {
            fc_pson_destroy(r->h);
            r->h = nullptr;
        }
        lk.lock();
        s->preparing--;
        if (rc == -6) {
            // the pool is fragmented: wait for a release, then try again
            s->committed -= r->need;
            if (s->committed > 0 || s->preparing > 0) {
                s->queue.push_front(r);
                s->cv.wait(lk, [s, rel0] { return s->stop || s->releases != rel0; });
                continue;
            }
        } else if (rc) {
            s->committed -= r->need;
        }
        if (rc) {
            r->rc = rc;
            r->err = err;
            s->failed.push_back(r);
        } else {
            s->ready.push_back(r);
        }
        s->cv.notify_all();
    }
}

// ready requests take free slots; admission errors are reported
int admit(fc_serve *s)
{
    std::vector<Request *> in, bad;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        while (!s->ready.empty() && s->free_slots.size() > in.size()) {
            in.push_back(s->ready.front());
            s->ready.pop_front();
        }
        while (!s->failed.empty()) {
            bad.push_back(s->failed.front());
            s->failed.pop_front();
        }
    }
    for (Request *r : bad) {
        set_error(r->rc, r->err);
        finish(s, r, r->rc);
    }
    for (Request *r : in) {
        r->slot = s->free_slots.back();
        s->free_slots.pop_back();
        s->h_desc[r->slot] = r->desc;
        SCHK(hipMemcpyAsync(s->d_desc + r->slot, s->h_desc + r->slot, sizeof(ReqDesc), hipMemcpyHostToDevice,
                            s->xs));
        r->joining = true;
        r->iter = 0;
        memset(&r->st, 0, sizeof(r->st));
        s->active.push_back(r);
    }
    return 0;
}

int run(fc_serve *s)
{
    if (s->ent.size() > (size_t)ENT_CAP) return set_error(-1, "batch program larger than the entry buffer");
    if (!s->ent.empty()) {
        memcpy(s->h_ent, s->ent.data(), sizeof(BEnt) * s->ent.size());
        SCHK(hipMemcpyAsync(s->d_ent, s->h_ent, sizeof(BEnt) * s->ent.size(), hipMemcpyHostToDevice, s->xs));
    }
    for (const Launch &L : s->launches) L.fn(s->d_ent + L.off, L.ne, L.nb);
    s->last_launches = (int)s->launches.size();
    return 0;
}

}  // namespace

// ===========================================================================
int fc_serve_create(const fc_serve_config *cfg, fc_serve **out)
{
    if (!cfg || !out) return set_error(-1, "null argument");
    if (cfg->max_active < 1 || cfg->max_active > 256) return set_error(-1, "max_active must be 1..256");
    if (cfg->pool_bytes < ((size_t)1 << 20)) return set_error(-1, "pool_bytes must be at least 1 MiB");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(-5, "no HIP device available");
    fc_serve *s = new fc_serve();
    s->cfg = *cfg;
    int rc = 0;
    if (hipStreamCreateWithFlags(&s->xs, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->as, hipStreamNonBlocking) != hipSuccess)
        rc = set_error(-3, "hipStreamCreate failed");
    if (!rc) rc = s->pool.init(cfg->pool_bytes);
    const int M = cfg->max_active;
    if (!rc && (hipMalloc((void **)&s->d_desc, sizeof(ReqDesc) * M) != hipSuccess ||
                hipHostMalloc((void **)&s->h_desc, sizeof(ReqDesc) * M, 0) != hipSuccess ||
                hipMalloc((void **)&s->d_ent, sizeof(BEnt) * ENT_CAP) != hipSuccess ||
                hipHostMalloc((void **)&s->h_ent, sizeof(BEnt) * ENT_CAP, 0) != hipSuccess ||
                hipMalloc((void **)&s->d_out, sizeof(double) * 4 * M) != hipSuccess ||
                hipHostMalloc((void **)&s->h_out, sizeof(double) * 4 * M, 0) != hipSuccess))
        rc = set_error(-4, "server buffer allocation failed");
    if (rc) {
        fc_serve_destroy(s);
        return rc;
    }
    for (int i = M - 1; i >= 0; --i) s->free_slots.push_back(i);
    const int nw = cfg->admit_threads > 0 ? std::min(cfg->admit_threads, 64) : 8;
    for (int w = 0; w < nw; ++w) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            fc_serve_destroy(s);
            return set_error(-3, "hipStreamCreate failed");
        }
        s->wst.push_back(st);
    }
    for (int w = 0; w < nw; ++w) s->workers.emplace_back(worker, s, w);
    *out = s;
    return 0;
}

int fc_serve_destroy(fc_serve *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    for (std::thread &t : s->workers) t.join();
    if (s->xs) hipStreamSynchronize(s->xs);
    if (s->as) hipStreamSynchronize(s->as);
    for (Request *r : s->active) { release(s, r); delete r; }
    for (Request *r : s->ready) { release(s, r); delete r; }
    for (Request *r : s->queue) delete r;
    for (Request *r : s->failed) delete r;
    for (hipStream_t st : s->wst) hipStreamDestroy(st);
    s->pool.release();
    if (s->d_desc) hipFree(s->d_desc);
    if (s->h_desc) hipHostFree(s->h_desc);
    if (s->d_ent) hipFree(s->d_ent);
    if (s->h_ent) hipHostFree(s->h_ent);
    if (s->d_out) hipFree(s->d_out);
    if (s->h_out) hipHostFree(s->h_out);
    if (s->xs) hipStreamDestroy(s->xs);
    if (s->as) hipStreamDestroy(s->as);
    delete s;
    return 0;
}

int fc_serve_submit(fc_serve *s, const fc_serve_request *rq, long long *id)
{
    if (!s || !rq) return set_error(-1, "null argument");
    if (!rq->grid.x || !rq->grid.y || !rq->grid.z || !rq->iblank || !rq->ghost || !rq->div || !rq->pPrime)
        return set_error(-1, "null request field");
    Request *r = new Request();
    r->id = s->next_id++;
    r->rq = *rq;
    r->t_submit = Clock::now();
    r->need = handle_footprint(&rq->grid, &rq->params, rq->nGhost, rq->nDead);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queue.push_back(r);
        s->cv.notify_all();
    }
    if (id) *id = r->id;
    return 0;
}

int fc_serve_step(fc_serve *s, fc_serve_status *status)
{
    if (!s) return set_error(-1, "null argument");
    int arc = admit(s);
    if (arc) return arc;
    if (s->active.empty()) {
        // nothing to run yet: wait (briefly) for a worker to finish a handle
        std::unique_lock<std::mutex> lk(s->mu);
        if (!s->queue.empty() || s->preparing > 0)
            s->cv.wait_for(lk, std::chrono::milliseconds(20), [s] { return !s->ready.empty() || !s->failed.empty(); });
        lk.unlock();
        if ((arc = admit(s))) return arc;
    }
    s->last_launches = 0;
    if (!s->active.empty()) {
        s->ent.clear();
        s->launches.clear();
        Builder B{s};
        const ReqDesc *D = s->d_desc;
        hipStream_t xs = s->xs;
        std::vector<Request *> &A = s->active, M, F;
        for (Request *r : A) (r->shape.mixed ? M : F).push_back(r);
        auto all = [](Request *) { return true; };
        auto rows = [](Request *r) { return dim3(r->shape.nrb); };
        // the mixing step that makes this iteration's iterate: the incoming
        // pressure (pinned) for a request that just joined, G - dG*gamma else
        B.phase(A, [](Request *r) { return r->joining; }, [](Request *) { return dim3(1); },
                [xs, D](const BEnt *E, int ne, int nb) { b_aa_lsq(xs, E, ne, nb, D, 1); });
        B.phase(A, [](Request *r) { return r->joining; }, rows,
                [xs, D](const BEnt *E, int ne, int nb) { b_aa_next(xs, E, ne, nb, D, 1); });
        B.phase(A, [](Request *r) { return !r->joining; }, rows,
                [xs, D](const BEnt *E, int ne, int nb) { b_aa_next(xs, E, ne, nb, D, 0); });
        // loop body: outer BC, GCM update in dependency order, V-cycle(s)
        B.phase(A, all, [](Request *r) { return dim3(cdiv(face_cells(r->shape.d[0]), 256)); },
                [xs, D](const BEnt *E, int ne, int nb) { b_bc_fill(xs, E, ne, nb, D, 0, 0, 1, 0); });
        int gmax = 0;
        for (Request *r : A) gmax = std::max(gmax, (int)r->shape.glevel.size() - 1);
        for (int l = 0; l < gmax; ++l)
            B.phase(A, [l](Request *r) { return l + 1 < (int)r->shape.glevel.size(); },
                    [l](Request *r) { return dim3(cdiv(r->shape.glevel[l + 1] - r->shape.glevel[l], 256)); },
                    [xs, D, l](const BEnt *E, int ne, int nb) { b_gcm_level(xs, E, ne, nb, D, l); });
        B.phase(M, all, rows, [xs, D](const BEnt *E, int ne, int nb) { b_resid64_to32(xs, E, ne, nb, D); });
        vcycle_phases(B, s, M, 1);
        vcycle_phases(B, s, F, 0);
        // mixed: the GCM changes join the FP32 correction (e = G - p)
        B.phase(M, [](Request *r) { return r->shape.glevel.back() > 0 && (int)r->shape.glevel.size() > 1; },
                [](Request *r) { return dim3(cdiv(r->shape.glevel.back(), 256)); },
                [xs, D](const BEnt *E, int ne, int nb) { b_gcm_fold(xs, E, ne, nb, D, 0); });
        // sums of the stop test, Anderson history and mixing coefficients
        B.phase(A, all, rows, [xs, D](const BEnt *E, int ne, int nb) { b_aa_resid(xs, E, ne, nb, D); });
        B.phase(A, all, [](Request *) { return dim3(ROW_COMPONENTS); },
                [xs, D](const BEnt *E, int ne, int nb) { b_reduce_cols(xs, E, ne, nb, D); });
        B.phase(A, all, [](Request *) { return dim3(1); },
                [xs, D](const BEnt *E, int ne, int nb) { b_aa_lsq(xs, E, ne, nb, D, 0); });
        double *dout = s->d_out;
        B.phase(A, all, [](Request *) { return dim3(1); },
                [xs, D, dout](const BEnt *E, int ne, int nb) { b_gather(xs, E, ne, nb, D, dout); });
        int rc = run(s);
        if (rc) return rc;
        SCHK(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * 4 * s->cfg.max_active, hipMemcpyDeviceToHost, xs));
        SCHK(hipStreamSynchronize(xs));
        s->steps++;
        // stop tests (as solve_poisson with outer_rule RELRES)
        std::vector<Request *> keep, fin;
        for (Request *r : A) {
            r->joining = false;
            r->iter++;
            const double *o = s->h_out + 4 * r->slot;
            const double s2 = o[0], d2 = o[1], maxf = o[2], maxg = o[3];
            const fc_pson_params &P = r->rq.params;
            const long nn = (long)r->rq.grid.nx * r->rq.grid.ny * r->rq.grid.nz;
            r->st.outer_iters = r->iter;
            r->st.maxres = std::sqrt(s2 / (double)nn);
            r->st.maxres2 = maxf;
            r->st.rel_resid = (d2 > 0.0) ? std::sqrt(s2 / d2) : std::sqrt(s2);
            r->st.gcm_change = maxf;
            r->st.gcm_sweeps += (int)r->shape.glevel.size() > 1 ? 1 : 0;
            if (r->shape.mixed) { r->st.vcycles_fp32++; r->st.refine_steps++; }
            else r->st.vcycles_fp64++;
            const bool conv = r->st.rel_resid <= P.rel_tol && maxf <= P.rel_tol * std::max(maxg, 1e-300);
            if (conv) r->st.converged = 1;
            if (conv || r->iter >= P.itermax_pson) fin.push_back(r);
            else keep.push_back(r);
        }
        if (!fin.empty()) {
            // the answer is in p: the map value G (FP64), or the last iterate
            // after its GCM update (mixed: the update moves back into p)
            s->ent.clear();
            s->launches.clear();
            B.phase(fin, [](Request *r) { return r->shape.mixed && (int)r->shape.glevel.size() > 1 &&
                                                 r->shape.glevel.back() > 0; },
                    [](Request *r) { return dim3(cdiv(r->shape.glevel.back(), 256)); },
                    [xs, D](const BEnt *E, int ne, int nb) { b_gcm_fold(xs, E, ne, nb, D, 1); });
            if ((rc = run(s))) return rc;
            SCHK(hipStreamSynchronize(xs));
            for (Request *r : fin) {
                int drc = fc_pson_download(r->h, r->rq.pPrime, nullptr);
                finish(s, r, drc ? drc : (r->st.converged ? 0 : 1));
            }
        }
        A.swap(keep);
    }
    if (status) {
        std::lock_guard<std::mutex> lk(s->mu);
        status->active = (int)s->active.size();
        status->queued = (int)(s->queue.size() + s->ready.size()) + s->preparing;
        status->finished = (int)(s->done.size() + s->failed.size());
        status->launches = s->last_launches;
        status->steps = s->steps;
        status->pool_used = s->pool.in_use();
        status->pool_capacity = s->pool.capacity();
    }
    return 0;
}

int fc_serve_poll(fc_serve *s, long long *id, int *rc, fc_pson_stats *stats)
{
    if (!s) return set_error(-1, "null argument");
    if (s->done.empty()) return 0;
    Done d = s->done.front();
    s->done.pop_front();
    if (id) *id = d.id;
    if (rc) *rc = d.rc;
    if (stats) *stats = d.st;
    return 1;
}
