// fc_serve.hip -- request server over the gfx950 pressure-Poisson path
// (include/fc_poisson.h, fc_serve_*): admission by device memory, resident
// hierarchies in one device pool, and continuous batching -- every step runs
// one accelerated outer iteration of solve_poisson (CORE_POISSON.F90:83-200
// with Anderson mixing) for all admitted requests, each phase of it as one
// batched launch (fc_kernels.hip, kb_*).  Requests join between steps and
// leave when they meet their own stop test.
//
// One host thread drives a server.  Setup of an admitted request (its handle:
// metrics, masks, GCM lists, uploads) runs on the admission stream; the
// batched phases run on the execution stream, which is idle at every step
// boundary (each step ends with the read-back of the stop-test sums).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <string>
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
    fc_pson *h = nullptr;
    int slot = -1;
    ReqShape shape;
    bool joining = false;
    int iter = 0;
    int rc = 0;
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
    hipStream_t xs = nullptr, as = nullptr;   // execution, admission
    Arena pool;
    ReqDesc *d_desc = nullptr, *h_desc = nullptr;
    BEnt *d_ent = nullptr, *h_ent = nullptr;
    double *d_out = nullptr, *h_out = nullptr;
    std::deque<Request *> queue;
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
void smooth_phases(Builder &B, fc_serve *s, const std::vector<Request *> &grp, int prec, int l,
                   const std::function<bool(Request *)> &take)
{
    const ReqDesc *D = s->d_desc;
    hipStream_t xs = s->xs;
    const int bcsel = (l == 0 && prec == 0) ? 0 : 1;   // FP64 level 1 carries the given outer BC
    B.phase(grp, take, [l](Request *r) { return dim3(cdiv(face_cells(r->shape.d[l]), 256)); },
            [xs, D, prec, l, bcsel](const BEnt *E, int ne, int nb) { b_bc_fill(xs, E, ne, nb, D, prec, l, 0, bcsel); });
    size_t lds = 0;
    for (Request *r : grp)
        if (take(r))
            for (int dir = 0; dir < 3; ++dir)
                lds = std::max(lds, (size_t)r->shape.sw[l][dir].NL * r->shape.sw[l][dir].Rp * 2 *
                                        (prec ? sizeof(float) : sizeof(double)));
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
        smooth_phases(B, s, grp, prec, l, [l](Request *r) { return r->shape.nlev > l; });
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
                [l](Request *r) {
                    const LevelDims &f = r->shape.d[l - 1];
                    return dim3(cdiv(f.nx - 1, 64), cdiv(f.ny - 1, 4), f.nz - 1);
                },
                [xs, D, prec, l](const BEnt *E, int ne, int nb) { b_prolong(xs, E, ne, nb, D, prec, l); });
    }
    smooth_phases(B, s, grp, prec, 0, [](Request *r) { return r->shape.cycle == 2; });
}

void release(fc_serve *s, Request *r)
{
    if (r->h) fc_pson_destroy(r->h);
    r->h = nullptr;
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

// create and describe the handle of a queued request; 1 = no room yet
int admit_one(fc_serve *s, Request *r)
{
    const fc_serve_request &q = r->rq;
    size_t need = handle_footprint(&q.grid, &q.params, q.nGhost, q.nDead);
    if (need > s->pool.capacity()) return set_error(-6, "request needs more device memory than the pool holds");
    if (need > s->pool.capacity() - s->pool.in_use()) return 1;
    fc_pson *h = nullptr;
    int rc = handle_create(&q.grid, &q.params, &s->pool, s->as, &h);
    if (rc == -6) return 1;     // fragmentation: wait for more memory
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
    r->slot = s->free_slots.back();
    s->free_slots.pop_back();
    if ((rc = handle_describe(h, &s->h_desc[r->slot], &r->shape))) return rc;
    SCHK(hipStreamSynchronize(s->as));
    SCHK(hipMemcpyAsync(s->d_desc + r->slot, s->h_desc + r->slot, sizeof(ReqDesc), hipMemcpyHostToDevice, s->xs));
    r->joining = true;
    r->iter = 0;
    memset(&r->st, 0, sizeof(r->st));
    return 0;
}

void admit(fc_serve *s)
{
    while (!s->queue.empty() && !s->free_slots.empty()) {
        Request *r = s->queue.front();
        int rc = admit_one(s, r);
        // no room yet (or the pool is too fragmented while others run): wait,
        // in arrival order; with nothing else resident it can never fit
        if (rc == 1 || rc == -6) {
            if (r->h) release(s, r);
            if (!s->active.empty()) break;
            rc = set_error(-6, "request does not fit the device memory pool");
        }
        s->queue.pop_front();
        if (rc) {
            finish(s, r, rc);
            continue;
        }
        s->active.push_back(r);
    }
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
    *out = s;
    return 0;
}

int fc_serve_destroy(fc_serve *s)
{
    if (!s) return 0;
    if (s->xs) hipStreamSynchronize(s->xs);
    if (s->as) hipStreamSynchronize(s->as);
    for (Request *r : s->active) { release(s, r); delete r; }
    for (Request *r : s->queue) delete r;
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
    s->queue.push_back(r);
    if (id) *id = r->id;
    return 0;
}

int fc_serve_step(fc_serve *s, fc_serve_status *status)
{
    if (!s) return set_error(-1, "null argument");
    admit(s);
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
        B.phase(M, all, rows, [xs, D](const BEnt *E, int ne, int nb) { b_add_into(xs, E, ne, nb, D, 0); });
        B.phase(M, all, [](Request *r) { return dim3(cdiv(face_cells(r->shape.d[0]), 256)); },
                [xs, D](const BEnt *E, int ne, int nb) { b_bc_fill(xs, E, ne, nb, D, 0, 0, 2, 0); });
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
            // the answer is the last map value G: p = G for mixed requests
            s->ent.clear();
            s->launches.clear();
            std::vector<Request *> fm;
            for (Request *r : fin) if (r->shape.mixed) fm.push_back(r);
            B.phase(fm, all, rows, [xs, D](const BEnt *E, int ne, int nb) { b_add_into(xs, E, ne, nb, D, 1); });
            B.phase(fm, all, [](Request *r) { return dim3(cdiv(face_cells(r->shape.d[0]), 256)); },
                    [xs, D](const BEnt *E, int ne, int nb) { b_bc_fill(xs, E, ne, nb, D, 0, 0, 1, 0); });
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
        status->active = (int)s->active.size();
        status->queued = (int)s->queue.size();
        status->finished = (int)s->done.size();
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
