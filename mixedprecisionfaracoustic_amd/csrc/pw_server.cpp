// pw_server.cpp -- the shot server of the C ABI (include/porowave.h).
//
// One worker thread owns the device.  Per round it
//   1. admits queued shots first come, first served while the slot count and
//      the device-memory budget allow (solver buffers, medium crop, source,
//      receivers, per-step tables);
//   2. advances every active shot by K steps, K = min(round_steps, fewest
//      steps any active shot has left), so shots finish on round boundaries:
//      per precision, one launch per step kernel covers all its shots (the
//      descriptor tables of pw_kernels.hip), and the K steps are captured once
//      as a hipGraph and replayed while the set of shots stays the same; or,
//      with cfg.persistent, one persistent cooperative launch per precision
//      runs the whole round;
//   3. waits for the round, copies the seismograms of finished shots to the
//      host and frees their device memory.
// Submitting, waiting and fetching results are thread-safe host operations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "pw_device.h"
#include "pw_solver.h"

using namespace pw;

namespace {

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Shot {
    long long id = 0;
    pw_shot req{};
    std::vector<int> rfield;
    std::vector<long long> rindex;
    int state = PW_SHOT_QUEUED;
    pw_status err = PW_OK;
    std::string errmsg;
    double t_submit = -1, t_admit = -1, t_done = -1;
    int waiters = 0;       // threads inside pw_server_wait on this shot (release refuses meanwhile)
    bool cancel = false;   // pw_server_cancel on a running shot: retired at the end of the round
    size_t bytes = 0;      // device bytes while admitted (estimate used for admission)
    long long cells = 0;
    // while admitted
    pw_solver* s = nullptr;
    void* d_tab = nullptr;  // amp [nt] (R), oscale [nt][2], ratio [nt]
    const void* amp = nullptr;
    const float* oscale = nullptr;
    const float* ratio = nullptr;
    int it = 0;
    // after completion
    std::vector<float> seis;
};

// Device bytes a shot holds while admitted: the solver pool, receivers and seismogram, the step tables.
size_t shot_bytes(const pw_shot& q)
{
    pw_config c;
    pw_config_default(&c);
    c.nx = q.nx; c.ny = q.ny; c.nz = q.nz; c.np = q.np;
    c.h = q.h; c.dt = q.dt; c.nt = q.nt; c.precision = q.precision;
    c.cpml_r = q.cpml_r; c.cpml_v0 = q.cpml_v0;
    size_t b = solver_pool_bytes(c);
    b += al256((size_t)q.nrec * 4) * 2 + al256((size_t)q.nrec * q.nt * 4);
    b += al256((size_t)q.nt * 8) + al256((size_t)q.nt * 8) + al256((size_t)q.nt * 4);
    return b;
}

}  // namespace

struct pw_server {
    pw_server_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::chrono::steady_clock::time_point t0;
    // medium volume (device, FP64, PW_NBASE dense fields)
    double* d_medium = nullptr;
    int mxi = 0, myi = 0, mzi = 0;
    size_t medium_bytes = 0;
    size_t budget = 0;
    bool auto_budget = false;
    // queue and records
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Shot*> pending;
    std::map<long long, std::unique_ptr<Shot>> shots;
    long long next_id = 1;
    bool stop = false;
    // worker state
    std::thread worker;
    std::vector<Shot*> active;
    int admitting = 0;  // shots being set up outside the lock
    size_t used = 0;
    pw_server_stats stats{};
    double clock() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// ---- admission: per-shot set-up on the worker thread ----
pw_status setup_shot(pw_server* srv, Shot* q)
{
    const pw_shot& r = q->req;
    pw_config c;
    pw_config_default(&c);
    c.nx = r.nx; c.ny = r.ny; c.nz = r.nz; c.np = r.np;
    c.h = r.h; c.dt = r.dt; c.nt = r.nt; c.precision = r.precision;
    c.cpml_r = r.cpml_r; c.cpml_v0 = r.cpml_v0;
    c.device = srv->cfg.device;
    c.stream = srv->stream;
    pw_status st = solver_create(&c, &q->s);
    if (st != PW_OK) return st;
    pw_solver* s = q->s;
    const int nxi = r.nx - 2 * r.np, nyi = r.ny - 2 * r.np, nzi = r.nz - 2 * r.np;
    const size_t ni = (size_t)srv->mxi * srv->myi * srv->mzi;
    const long long rs = srv->mxi, ps = (long long)srv->mxi * srv->myi;
    const long long o0 = (long long)r.crop[0] + (long long)r.crop[1] * rs + (long long)r.crop[2] * ps;
    const double* view[PW_NBASE];
    for (int k = 0; k < PW_NBASE; ++k) view[k] = srv->d_medium + k * ni + o0;
    if ((st = install_medium(s, view, nxi, nyi, nzi, rs, ps)) != PW_OK) return st;
    if ((st = pw_set_source(s, r.sx, r.sy, r.sz, r.f0, r.w)) != PW_OK) return st;
    if ((st = pw_set_receivers(s, r.nrec, q->rfield.data(), q->rindex.data())) != PW_OK) return st;
    // per-step tables, planned exactly as pw_run plans its steps
    const int nt = r.nt;
    const bool f64 = r.precision == PW_FP64;
    const size_t o_os = al256((size_t)nt * (f64 ? 8 : 4)), o_ra = o_os + al256((size_t)nt * 8);
    const size_t total = o_ra + al256((size_t)nt * 4);
    std::vector<char> host(total, 0);
    float dyn = 1.0f;
    for (int it = 0; it < nt; ++it) {
        const StepPlan p = plan_step(s, it, &dyn);
        if (f64) ((double*)host.data())[it] = p.amp;
        else ((float*)host.data())[it] = (float)p.amp;
        float* os = (float*)(host.data() + o_os);
        os[2 * it] = p.s_stress;
        os[2 * it + 1] = p.s_vel;
        ((float*)(host.data() + o_ra))[it] = p.ratio;
    }
    PW_HIP_TRY(hipMalloc(&q->d_tab, total));
    PW_HIP_TRY(hipMemcpyAsync(q->d_tab, host.data(), total, hipMemcpyHostToDevice, srv->stream));
    PW_HIP_TRY(hipStreamSynchronize(srv->stream));
    q->amp = q->d_tab;
    q->oscale = (const float*)((char*)q->d_tab + o_os);
    q->ratio = is_scaled(r.precision) ? (const float*)((char*)q->d_tab + o_ra) : nullptr;
    q->it = 0;
    return PW_OK;
}

void teardown_shot(Shot* q)
{
    if (q->s) pw_destroy(q->s);
    q->s = nullptr;
    if (q->d_tab) (void)hipFree(q->d_tab);
    q->d_tab = nullptr;
}

// ---- the rounds ----
struct Engine {
    pw_server* srv;
    BatchGroup groups[5];         // one per precision
    std::vector<Shot*> members;   // active shots in group order when the tables were built
    bool stale = true;            // membership changed since the tables / graph were built
    hipGraphExec_t exec = nullptr;
    int exec_K = 0;
    int last_K = 0;               // length of the previous round with the same shots
    long long launches = 0, graphs_built = 0;  // added to the stats under the lock
    double graph_s = 0.0;

    ~Engine()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (BatchGroup& g : groups) batch_free(g);
    }

    pw_status rebuild()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        last_K = 0;
        members.clear();
        for (int p = 0; p < 5; ++p) {
            std::vector<BatchShot> bs;
            std::vector<int> its;
            for (Shot* q : srv->active) {
                if (q->req.precision != p) continue;
                BatchShot b{};
                b.b = &q->s->b;
                b.amp = q->amp;
                b.oscale = q->oscale;
                b.ratio = q->ratio;
                b.rfield = q->s->d_rfield;
                b.roff = q->s->d_roff;
                b.seis = q->s->d_seis;
                b.nrec = q->s->nrec;
                b.nt = q->req.nt;
                b.src_off = q->s->src_off;
                b.has_src = q->s->has_source ? 1 : 0;
                for (int k = 0; k < 4; ++k) b.w[k] = q->s->w[k];
                bs.push_back(b);
                its.push_back(q->it);
                members.push_back(q);
            }
            PW_HIP_TRY(batch_build(groups[p], p, bs.data(), (int)bs.size(), its.data(), srv->stream));
        }
        stale = false;
        return PW_OK;
    }

    pw_status issue(int K)
    {
        for (int k = 0; k < K; ++k)
            for (const BatchGroup& g : groups) {
                if (g.n <= 0) continue;
                PW_HIP_TRY(batch_step(g, srv->stream, k));
                launches += (is_scaled(g.prec) ? 5 : 4) - (g.nrec > 0 ? 0 : 1);
            }
        for (const BatchGroup& g : groups)
            if (g.n > 0) {
                PW_HIP_TRY(batch_advance(g, srv->stream, K));
                launches += 1;
            }
        return PW_OK;
    }

    pw_status round(int K)
    {
        const pw_server_config& c = srv->cfg;
        if (!c.batched) {
            for (Shot* q : srv->active) {
                pw_status st = pw_run(q->s, K);
                if (st != PW_OK) return st;
                launches += 4L * K;
            }
            return PW_OK;
        }
        if (stale) {
            pw_status st = rebuild();
            if (st != PW_OK) return st;
        }
        if (c.persistent) {
            // one cooperative launch per precision for the whole round
            for (const BatchGroup& g : groups) {
                if (g.n <= 0) continue;
                PW_HIP_TRY(batch_round(g, srv->stream, K));
                PW_HIP_TRY(batch_advance(g, srv->stream, K));
                launches += 2;
            }
            return PW_OK;
        }
        // A graph pays off only when it is replayed: capture a round when the
        // previous round ran the same shots for the same number of steps.
        const bool repeat = last_K == K;
        last_K = K;
        if (!c.use_graphs || (!repeat && !(exec && exec_K == K))) return issue(K);
        if (!exec || exec_K != K) {
            const double tg = srv->clock();
            if (exec) (void)hipGraphExecDestroy(exec);
            exec = nullptr;
            PW_HIP_TRY(hipStreamBeginCapture(srv->stream, hipStreamCaptureModeThreadLocal));
            const long long l0 = launches;
            pw_status st = issue(K);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(srv->stream, &graph);
            launches = l0;
            if (st != PW_OK) {
                if (graph) (void)hipGraphDestroy(graph);
                return st;
            }
            if (e != hipSuccess) return fail(PW_ERR_HIP, std::string("round capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                exec = nullptr;
                return fail(PW_ERR_HIP, std::string("round instantiate: ") + hipGetErrorString(e));
            }
            exec_K = K;
            graphs_built++;
            graph_s += srv->clock() - tg;
        }
        PW_HIP_TRY(hipGraphLaunch(exec, srv->stream));
        launches += 1;
        return PW_OK;
    }
};

void finish(pw_server* srv, Shot* q, pw_status st, const std::string& msg)
{
    // called with srv->mu held
    q->t_done = srv->clock();
    if (st == PW_OK) {
        q->state = PW_SHOT_DONE;
        srv->stats.completed++;
        srv->stats.cell_steps += (double)q->cells * q->req.nt;
    } else {
        q->state = PW_SHOT_FAILED;
        q->err = st;
        q->errmsg = msg;
        srv->stats.failed++;
    }
    srv->cv_done.notify_all();
}

// The queued shot admission considers next: the front (first come first
// served), or with max_wait_s > 0 the oldest shot queued longer than that, else
// the one with the least work (cells x steps; ties: the earlier one).
std::deque<Shot*>::iterator next_pending(pw_server* srv)
{
    auto& pq = srv->pending;
    if (srv->cfg.max_wait_s <= 0.0) return pq.begin();
    const double now = srv->clock();
    auto best = pq.begin();
    for (auto it = pq.begin(); it != pq.end(); ++it) {  // queue order is arrival order
        if (now - (*it)->t_submit > srv->cfg.max_wait_s) return it;
        const double w = (double)(*it)->cells * (*it)->req.nt, wb = (double)(*best)->cells * (*best)->req.nt;
        if (w < wb) best = it;
    }
    return best;
}

void worker_main(pw_server* srv)
{
    (void)hipSetDevice(srv->cfg.device);
    Engine eng{srv};
    std::unique_lock<std::mutex> lk(srv->mu);
    for (;;) {
        // 1. admission, first come first served
        while (!srv->stop && !srv->pending.empty() && (int)srv->active.size() < srv->cfg.max_active) {
            const auto qi = next_pending(srv);
            Shot* q = *qi;
            if (q->bytes > srv->budget) {
                srv->pending.erase(qi);
                finish(srv, q, PW_ERR_OUT_OF_MEMORY,
                       "shot needs " + std::to_string(q->bytes) + " device bytes, above the server budget of " +
                           std::to_string(srv->budget));
                continue;
            }
            if (srv->used + q->bytes > srv->budget) break;
            // grid cells already admitted: a shot that would take them past
            // fill_cells waits for a slot (unless nothing is running)
            if (srv->cfg.fill_cells > 0 && !srv->active.empty()) {
                long long held = 0;
                for (const Shot* a : srv->active) held += a->cells;
                if (held + q->cells > srv->cfg.fill_cells) break;
            }
            srv->pending.erase(qi);
            q->state = PW_SHOT_RUNNING;
            q->t_admit = srv->clock();
            srv->admitting++;
            lk.unlock();
            const double ts = srv->clock();
            pw_status st = setup_shot(srv, q);
            const std::string msg = st == PW_OK ? std::string() : std::string(pw_last_error());
            const double te = srv->clock();
            lk.lock();
            srv->stats.setup_s += te - ts;
            srv->admitting--;
            if (st != PW_OK) {
                teardown_shot(q);
                finish(srv, q, st, msg);
                continue;
            }
            srv->active.push_back(q);
            srv->used += q->bytes;
            srv->stats.peak_bytes = std::max(srv->stats.peak_bytes, (long long)srv->used);
            srv->stats.peak_active = std::max(srv->stats.peak_active, (int)srv->active.size());
            eng.stale = true;
        }
        if (srv->active.empty()) {
            if (srv->stop) break;
            srv->cv_work.wait(lk);
            continue;
        }
        // 2. one round
        int K = srv->cfg.round_steps;
        for (Shot* q : srv->active) K = std::min(K, q->req.nt - q->it);
        lk.unlock();
        const double tr = srv->clock();
        pw_status st = eng.round(K);
        std::string msg = st == PW_OK ? std::string() : std::string(pw_last_error());
        if (st == PW_OK) {
            hipError_t e = hipStreamSynchronize(srv->stream);
            if (e != hipSuccess) {
                st = PW_ERR_HIP;
                msg = std::string("round: ") + hipGetErrorString(e);
            }
        }
        // 3. retire finished shots
        const double tw = srv->clock();
        std::vector<Shot*> done;
        // each retired shot's own status: a failed read-back fails that shot only
        std::vector<std::pair<pw_status, std::string>> outcome;
        if (st == PW_OK) {
            std::vector<char> cancelled;
            {
                // cancel flags are set under the lock by pw_server_cancel
                std::lock_guard<std::mutex> lc(srv->mu);
                for (Shot* q : srv->active) {
                    q->it += K;
                    if (q->it >= q->req.nt || q->cancel) {
                        done.push_back(q);
                        cancelled.push_back(q->it < q->req.nt ? 1 : 0);
                    }
                }
            }
            for (size_t i = 0; i < done.size(); ++i) {
                Shot* q = done[i];
                if (cancelled[i]) {
                    outcome.emplace_back(PW_ERR_CANCELLED, "shot cancelled after " + std::to_string(q->it) + " steps");
                    continue;
                }
                q->seis.assign((size_t)q->s->nrec * q->req.nt, 0.0f);
                pw_status g = q->s->nrec > 0 ? pw_get_seismogram(q->s, q->seis.data()) : PW_OK;
                std::string m = g == PW_OK ? std::string() : std::string(pw_last_error());
                if (g == PW_OK) {
                    // an unstable run (e.g. a source far beyond the float range)
                    // shows up as NaN or Inf at the receivers
                    for (float v : q->seis)
                        if (!std::isfinite(v)) {
                            g = PW_ERR_NONFINITE;
                            m = "shot's seismogram holds NaN or Inf values";
                            break;
                        }
                }
                outcome.emplace_back(g, m);
            }
        } else {
            done = srv->active;  // a failed round fails every shot in it
            outcome.assign(done.size(), {st, msg});
        }
        for (Shot* q : done) teardown_shot(q);
        const double tend = srv->clock();
        lk.lock();
        srv->stats.rounds++;
        srv->stats.busy_s += tend - tr;
        srv->stats.retire_s += done.empty() ? 0.0 : tend - tw;
        srv->stats.launches += eng.launches;
        srv->stats.graphs_built += eng.graphs_built;
        srv->stats.graph_s += eng.graph_s;
        eng.launches = eng.graphs_built = 0;
        eng.graph_s = 0.0;
        for (size_t i = 0; i < done.size(); ++i) {
            Shot* q = done[i];
            finish(srv, q, outcome[i].first, outcome[i].second);
            srv->used -= q->bytes;
            srv->active.erase(std::find(srv->active.begin(), srv->active.end(), q));
        }
        if (!done.empty()) eng.stale = true;
    }
    // fail whatever is still queued
    while (!srv->pending.empty()) {
        Shot* q = srv->pending.front();
        srv->pending.pop_front();
        finish(srv, q, PW_ERR_STATE, "server destroyed before the shot was admitted");
    }
}

pw_status check_shot(const pw_server* srv, const pw_shot* r)
{
    if (!r) return fail(PW_ERR_INVALID, "pw_server_submit: null shot");
    if (!srv->d_medium) return fail(PW_ERR_STATE, "pw_server_submit: no medium set");
    if (r->precision < PW_FP32 || r->precision > PW_COMP16F) return fail(PW_ERR_INVALID, "pw_server_submit: precision");
    if (r->np < 1 || r->nx < 2 * r->np + 3 || r->ny < 2 * r->np + 3 || r->nz < 2 * r->np + 3)
        return fail(PW_ERR_INVALID, "pw_server_submit: grid must be at least 2*np+3 cells per axis");
    if (!(r->h > 0) || !(r->dt > 0) || r->nt < 1 || !(r->f0 > 0))
        return fail(PW_ERR_INVALID, "pw_server_submit: h, dt, nt and f0 must be positive");
    const int ni[3] = {r->nx - 2 * r->np, r->ny - 2 * r->np, r->nz - 2 * r->np};
    const int mi[3] = {srv->mxi, srv->myi, srv->mzi};
    for (int a = 0; a < 3; ++a)
        if (r->crop[a] < 0 || r->crop[a] + ni[a] > mi[a])
            return fail(PW_ERR_INVALID, "pw_server_submit: crop outside the medium volume on axis " + std::to_string(a));
    if (r->sx < 1 || r->sy < 1 || r->sz < 1 || r->sx > r->nx - 2 || r->sy > r->ny - 2 || r->sz > r->nz - 2)
        return fail(PW_ERR_INVALID, "pw_server_submit: source cell must be inside the updated grid");
    if (r->nrec < 0 || (r->nrec > 0 && (!r->rec_field || !r->rec_index)))
        return fail(PW_ERR_INVALID, "pw_server_submit: receivers");
    const long long ncell = (long long)r->nx * r->ny * r->nz;
    for (int k = 0; k < r->nrec; ++k)
        if (r->rec_field[k] < 0 || r->rec_field[k] >= PW_NFIELDS || r->rec_index[k] < 0 || r->rec_index[k] >= ncell)
            return fail(PW_ERR_INVALID, "pw_server_submit: receiver " + std::to_string(k) + " out of range");
    return PW_OK;
}

Shot* find_shot(pw_server* srv, long long id)
{
    auto it = srv->shots.find(id);
    return it == srv->shots.end() ? nullptr : it->second.get();
}

pw_status put_medium(pw_server* srv, double* d_base, int nxi, int nyi, int nzi)
{
    if (srv->d_medium) (void)hipFree(srv->d_medium);
    srv->d_medium = d_base;
    srv->mxi = nxi; srv->myi = nyi; srv->mzi = nzi;
    const size_t bytes = (size_t)PW_NBASE * nxi * nyi * nzi * sizeof(double);
    if (srv->auto_budget) {
        srv->budget += srv->medium_bytes;
        srv->budget = srv->budget > bytes ? srv->budget - bytes : 0;
    }
    srv->medium_bytes = bytes;
    return PW_OK;
}

bool idle(pw_server* srv) { return srv->pending.empty() && srv->active.empty() && srv->admitting == 0; }

}  // namespace

extern "C" void pw_server_config_default(pw_server_config* c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->stream = nullptr;
    c->hbm_budget = 0;
    c->max_active = 16;
    c->round_steps = 32;
    c->use_graphs = 1;
    c->batched = 1;
    c->fill_cells = 1LL << 23;
    c->persistent = 0;
    c->max_wait_s = 0.0;
}

extern "C" pw_status pw_server_create(const pw_server_config* cfg, pw_server** out)
{
    if (!cfg || !out) return fail(PW_ERR_INVALID, "pw_server_create: null argument");
    *out = nullptr;
    if (cfg->max_active < 1 || cfg->round_steps < 1 || cfg->hbm_budget < 0 || cfg->fill_cells < 0)
        return fail(PW_ERR_INVALID,
                    "pw_server_create: max_active and round_steps must be >= 1, budget and fill_cells >= 0");
    if (cfg->persistent != 0 && cfg->persistent != 1)
        return fail(PW_ERR_INVALID, "pw_server_create: persistent must be 0 or 1");
    if (cfg->persistent && !cfg->batched)
        return fail(PW_ERR_INVALID, "pw_server_create: persistent rounds need batched rounds");
    if (!(cfg->max_wait_s >= 0.0) || !std::isfinite(cfg->max_wait_s))
        return fail(PW_ERR_INVALID, "pw_server_create: max_wait_s must be finite and >= 0");
    int ndev = 0;
    PW_HIP_TRY(hipGetDeviceCount(&ndev));
    if (cfg->device < 0 || cfg->device >= ndev) return fail(PW_ERR_INVALID, "pw_server_create: no such device");
    PW_HIP_TRY(hipSetDevice(cfg->device));
    if (cfg->persistent) {
        int coop = 0;
        PW_HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, cfg->device));
        if (!coop) return fail(PW_ERR_UNSUPPORTED, "pw_server_create: the device has no cooperative launch");
    }
    std::unique_ptr<pw_server> srv(new pw_server());
    srv->cfg = *cfg;
    srv->t0 = std::chrono::steady_clock::now();
    if (cfg->hbm_budget > 0) {
        srv->budget = (size_t)cfg->hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        PW_HIP_TRY(hipMemGetInfo(&fr, &tot));
        srv->budget = (size_t)(0.9 * (double)fr);
        srv->auto_budget = true;
    }
    if (cfg->stream) {
        srv->stream = (hipStream_t)cfg->stream;
    } else {
        PW_HIP_TRY(hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking));
        srv->own_stream = true;
    }
    pw_server* p = srv.release();
    p->worker = std::thread(worker_main, p);
    *out = p;
    return PW_OK;
}

extern "C" void pw_server_destroy(pw_server* srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->cfg.device);
    if (srv->d_medium) (void)hipFree(srv->d_medium);
    if (srv->own_stream && srv->stream) (void)hipStreamDestroy(srv->stream);
    delete srv;
}

extern "C" pw_status pw_server_set_medium(pw_server* srv, const double* const* base, int nxi, int nyi, int nzi)
{
    if (!srv || !base) return fail(PW_ERR_INVALID, "pw_server_set_medium: null argument");
    if (nxi < 1 || nyi < 1 || nzi < 1) return fail(PW_ERR_INVALID, "pw_server_set_medium: empty volume");
    for (int k = 0; k < PW_NBASE; ++k)
        if (!base[k]) return fail(PW_ERR_INVALID, "pw_server_set_medium: null base field");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!idle(srv)) return fail(PW_ERR_STATE, "pw_server_set_medium: shots are queued or running");
    PW_HIP_TRY(hipSetDevice(srv->cfg.device));
    const size_t ni = (size_t)nxi * nyi * nzi;
    double* d = nullptr;
    PW_HIP_TRY(hipMalloc(&d, PW_NBASE * ni * sizeof(double)));
    hipError_t e = hipSuccess;
    for (int k = 0; k < PW_NBASE && e == hipSuccess; ++k)
        e = hipMemcpy(d + k * ni, base[k], ni * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(PW_ERR_HIP, std::string("pw_server_set_medium: upload: ") + hipGetErrorString(e));
    }
    return put_medium(srv, d, nxi, nyi, nzi);
}

extern "C" pw_status pw_server_set_medium_layered(pw_server* srv, const pw_layered* spec, int nxi, int nyi, int nzi,
                                                  int np)
{
    if (!srv || !spec) return fail(PW_ERR_INVALID, "pw_server_set_medium_layered: null argument");
    if (nxi < 1 || nyi < 1 || nzi < 1 || np < 0) return fail(PW_ERR_INVALID, "pw_server_set_medium_layered: sizes");
    if (spec->nlayers < 1 || spec->nlayers > PW_MAX_LAYERS || spec->z_top[0] != 0)
        return fail(PW_ERR_INVALID, "pw_server_set_medium_layered: layers");
    for (int k = 1; k < spec->nlayers; ++k)
        if (spec->z_top[k] <= spec->z_top[k - 1])
            return fail(PW_ERR_INVALID, "pw_server_set_medium_layered: z_top not ascending");
    if (spec->perturb < 0 || spec->perturb * (spec->clamp > 0 ? spec->clamp : 3.0) >= 1.0)
        return fail(PW_ERR_INVALID, "pw_server_set_medium_layered: perturb*clamp must be in [0, 1)");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (!idle(srv)) return fail(PW_ERR_STATE, "pw_server_set_medium_layered: shots are queued or running");
    PW_HIP_TRY(hipSetDevice(srv->cfg.device));
    const size_t ni = (size_t)nxi * nyi * nzi;
    double *d = nullptr, *w = nullptr, *part = nullptr;
    PW_HIP_TRY(hipMalloc(&d, PW_NBASE * ni * sizeof(double)));
    hipError_t e = hipMalloc(&w, 4 * ni * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&part, 2 * 4096 * sizeof(double));
    if (e == hipSuccess) e = generate_layered(*spec, nxi, nyi, nzi, np, d, w, part, srv->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(srv->stream);
    if (w) (void)hipFree(w);
    if (part) (void)hipFree(part);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(e == hipErrorOutOfMemory ? PW_ERR_OUT_OF_MEMORY : PW_ERR_HIP,
                    std::string("pw_server_set_medium_layered: ") + hipGetErrorString(e));
    }
    return put_medium(srv, d, nxi, nyi, nzi);
}

extern "C" pw_status pw_server_submit(pw_server* srv, const pw_shot* shot, long long* id)
{
    if (!srv || !id) return fail(PW_ERR_INVALID, "pw_server_submit: null argument");
    if (!shot) return fail(PW_ERR_INVALID, "pw_server_submit: null shot");
    if (shot->nrec < 0 || (shot->nrec > 0 && (!shot->rec_field || !shot->rec_index)))
        return fail(PW_ERR_INVALID, "pw_server_submit: receivers");
    // the record is built first (host copies, byte estimate); validation against
    // the medium and queueing then happen under one lock, so that
    // pw_server_set_medium cannot replace the medium between the two
    std::unique_ptr<Shot> q(new Shot());
    q->req = *shot;
    q->rfield.assign(shot->rec_field, shot->rec_field + shot->nrec);
    q->rindex.assign(shot->rec_index, shot->rec_index + shot->nrec);
    q->req.rec_field = nullptr;
    q->req.rec_index = nullptr;
    q->cells = (long long)shot->nx * shot->ny * shot->nz;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        pw_status st = check_shot(srv, shot);
        if (st != PW_OK) return st;
        if (srv->stop) return fail(PW_ERR_STATE, "pw_server_submit: server is stopping");
        q->bytes = shot_bytes(*shot);
        q->id = srv->next_id++;
        q->t_submit = srv->clock();
        *id = q->id;
        srv->pending.push_back(q.get());
        srv->shots[q->id] = std::move(q);
        srv->stats.submitted++;
    }
    srv->cv_work.notify_one();
    return PW_OK;
}

extern "C" pw_status pw_server_wait(pw_server* srv, long long id, double timeout_s)
{
    if (!srv) return fail(PW_ERR_INVALID, "pw_server_wait: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    Shot* q = find_shot(srv, id);
    if (!q) return fail(PW_ERR_INVALID, "pw_server_wait: unknown shot id");
    // q stays valid while the lock is released inside the waits: pw_server_release
    // refuses a shot with waiters
    q->waiters++;
    auto ready = [&] { return q->state == PW_SHOT_DONE || q->state == PW_SHOT_FAILED; };
    bool ok = true;
    if (timeout_s < 0) srv->cv_done.wait(lk, ready);
    else ok = srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready);
    q->waiters--;
    if (!ok) return fail(PW_ERR_TIMEOUT, "pw_server_wait: shot " + std::to_string(id) + " not finished");
    if (q->state == PW_SHOT_FAILED) return fail(q->err, q->errmsg);
    return PW_OK;
}

extern "C" pw_status pw_server_status(pw_server* srv, long long id, pw_shot_status* out)
{
    if (!srv || !out) return fail(PW_ERR_INVALID, "pw_server_status: null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    Shot* q = find_shot(srv, id);
    if (!q) return fail(PW_ERR_INVALID, "pw_server_status: unknown shot id");
    std::memset(out, 0, sizeof(*out));
    out->state = q->state;
    out->precision = q->req.precision;
    out->nt = q->req.nt;
    out->nrec = q->req.nrec;
    out->cells = q->cells;
    out->device_bytes = (long long)q->bytes;
    out->t_submit = q->t_submit;
    out->t_admit = q->t_admit;
    out->t_done = q->t_done;
    out->error = q->err;
    std::strncpy(out->error_text, q->errmsg.c_str(), sizeof(out->error_text) - 1);
    return PW_OK;
}

extern "C" pw_status pw_server_result(pw_server* srv, long long id, float* seis)
{
    if (!srv || !seis) return fail(PW_ERR_INVALID, "pw_server_result: null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    Shot* q = find_shot(srv, id);
    if (!q) return fail(PW_ERR_INVALID, "pw_server_result: unknown shot id");
    if (q->state == PW_SHOT_FAILED) return fail(q->err, q->errmsg);
    if (q->state != PW_SHOT_DONE) return fail(PW_ERR_STATE, "pw_server_result: shot not finished");
    std::memcpy(seis, q->seis.data(), q->seis.size() * sizeof(float));
    return PW_OK;
}

extern "C" pw_status pw_server_release(pw_server* srv, long long id)
{
    if (!srv) return fail(PW_ERR_INVALID, "pw_server_release: null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    Shot* q = find_shot(srv, id);
    if (!q) return fail(PW_ERR_INVALID, "pw_server_release: unknown shot id");
    if (q->state != PW_SHOT_DONE && q->state != PW_SHOT_FAILED)
        return fail(PW_ERR_STATE, "pw_server_release: shot still queued or running");
    if (q->waiters > 0) return fail(PW_ERR_STATE, "pw_server_release: another thread is waiting on the shot");
    srv->shots.erase(id);
    return PW_OK;
}

extern "C" pw_status pw_server_cancel(pw_server* srv, long long id)
{
    if (!srv) return fail(PW_ERR_INVALID, "pw_server_cancel: null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    Shot* q = find_shot(srv, id);
    if (!q) return fail(PW_ERR_INVALID, "pw_server_cancel: unknown shot id");
    if (q->state == PW_SHOT_DONE || q->state == PW_SHOT_FAILED)
        return fail(PW_ERR_STATE, "pw_server_cancel: shot already finished");
    if (q->state == PW_SHOT_QUEUED) {
        auto it = std::find(srv->pending.begin(), srv->pending.end(), q);
        if (it != srv->pending.end()) srv->pending.erase(it);
        finish(srv, q, PW_ERR_CANCELLED, "shot cancelled before admission");
        return PW_OK;
    }
    q->cancel = true;  // running (or being admitted): the worker retires it after the round
    return PW_OK;
}

extern "C" pw_status pw_server_get_stats(pw_server* srv, pw_server_stats* out)
{
    if (!srv || !out) return fail(PW_ERR_INVALID, "pw_server_get_stats: null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    return PW_OK;
}

extern "C" double pw_server_clock(pw_server* srv) { return srv ? srv->clock() : -1.0; }
