// gf_serve.hip -- request server for reconstruction jobs on one GPU.
//
// Continuous batching: requests wait in a FIFO and are admitted (first fit) as
// soon as a slot and their HBM footprint are free.  An admitted job is prepared
// by a worker thread on its own stream (upload, gridding, iteration set-up),
// then joins the group of iterating jobs of its grid shape.  The scheduler
// thread advances every group one iteration at a time: the transform passes of
// all jobs of a half-spectrum group go out as one batched launch per pass
// (launch row = job), jobs join and leave between iterations, and at most
// `lookahead` iterations of a group are queued on its stream ahead of the GPU.
// A finished job is cropped and read back by a worker, its device memory is
// freed, and the next requests are admitted.  Each job's arithmetic is the same
// as alone: batching only shares launches.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <tuple>

#include "gf_job.h"

using namespace gf;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct genfire_request {
    int64_t id = 0;
    genfire_params prm{};
    int dim1 = 0, dim2 = 0, P = 0;
    std::vector<float> proj;
    std::vector<double> ang;
    std::vector<uint8_t> sup;
    bool has_sup = false;
    int64_t footprint = 0;
    int state = GENFIRE_REQ_QUEUED;
    int status = GENFIRE_OK;
    std::string err;
    std::unique_ptr<genfire_job> job;
    int it = 0;                      // next iteration while ITERATING
    bool joined = false;             // group stream has waited on ev_ready
    hipEvent_t ev_ready = nullptr;   // end of preparation on the job's stream
    hipEvent_t ev_done = nullptr;    // end of the last iteration on the group stream
    std::tuple<int, int, int, int> group{};
    // results
    std::vector<float> recon, errK, errR;
    double t_submit = 0, t_admit = 0, t_ready = 0, t_done = 0;

    ~genfire_request()
    {
        if (ev_ready) (void)hipEventDestroy(ev_ready);
        if (ev_done) (void)hipEventDestroy(ev_done);
    }
};

struct genfire_server {
    genfire_ctx* ctx = nullptr;
    genfire_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_task;
    std::deque<genfire_request*> queue;
    std::map<int64_t, std::unique_ptr<genfire_request>> reqs;
    std::vector<genfire_request*> active;
    std::deque<std::function<void()>> tasks;
    int64_t used = 0, peak_used = 0, next_id = 1;
    bool stop = false, mem_blocked = false;
    // GENFIRE_SERVER_FAIL_AT=k (tests of the failure path): the scheduler's k-th
    // group iteration raises a HIP-class error after its launches are queued, once
    int64_t fail_at = 0;
    std::thread sched;
    std::vector<std::thread> workers;

    struct Group {
        hipStream_t st = nullptr;
        std::deque<hipEvent_t> inflight, pool;
        // the batched steady-state iteration of the requests gids, captured once
        // and replayed while the group's membership stays the same
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        std::vector<int64_t> gids;
        void drop_graph()
        {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
            gids.clear();
        }
    };
    std::map<std::tuple<int, int, int, int>, Group> groups;   // touched by the scheduler thread only

    // statistics
    std::vector<double> lat;
    double t_first = 0, t_last = 0;
    int64_t submitted = 0, completed = 0, failed = 0, batched_launches = 0, batched_jobs = 0, graph_replays = 0;
    bool graphs = true;   // GENFIRE_GRAPHS=0 launches every iteration directly
    int max_active_seen = 0;

    // queues a worker task; the caller holds mu
    void post_locked(std::function<void()> f)
    {
        tasks.push_back(std::move(f));
        cv_task.notify_one();
    }
    void worker_loop();
    void sched_loop();
    bool admit_locked();
    void prepare(genfire_request* r);
    void finish(genfire_request* r);
    void fail(genfire_request* r, int code, const std::string& msg);
    bool advance_groups();
};

void genfire_server::worker_loop()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        std::function<void()> f;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_task.wait(lk, [&] { return stop || !tasks.empty(); });
            if (tasks.empty()) return;
            f = std::move(tasks.front());
            tasks.pop_front();
        }
        f();
    }
}

// First fit over the FIFO: the oldest request whose footprint fits is admitted.
// The caller holds mu.
bool genfire_server::admit_locked()
{
    bool any = false;
    if (mem_blocked) return false;
    for (auto q = queue.begin(); q != queue.end();) {
        genfire_request* r = *q;
        if ((int)active.size() >= cfg.max_active) break;
        if (used + r->footprint > cfg.hbm_budget_bytes) { ++q; continue; }
        q = queue.erase(q);
        used += r->footprint;
        peak_used = std::max(peak_used, used);
        r->state = GENFIRE_REQ_PREPARING;
        r->t_admit = now_s();
        active.push_back(r);
        max_active_seen = std::max(max_active_seen, (int)active.size());
        post_locked([this, r] { prepare(r); });
        any = true;
    }
    return any;
}

void genfire_server::fail(genfire_request* r, int code, const std::string& msg)
{
    r->job.reset();
    std::lock_guard<std::mutex> lk(mu);
    r->status = code;
    r->err = msg;
    r->state = GENFIRE_REQ_FAILED;
    r->t_done = now_s();
    used -= r->footprint;
    active.erase(std::remove(active.begin(), active.end(), r), active.end());
    ++failed;
    mem_blocked = false;
    cv_done.notify_all();
    cv_sched.notify_one();
}

void genfire_server::prepare(genfire_request* r)
{
    try {
        GF_HIP(hipSetDevice(ctx->device));
        auto j = std::make_unique<genfire_job>();
        job_setup(j.get(), ctx, &r->prm, r->dim1, r->dim2, r->P);
        job_set_inputs(j.get(), r->proj.data(), r->ang.data(), r->has_sup ? r->sup.data() : nullptr, nullptr);
        do_grid(j.get());
        release_grid_scratch(j.get());   // before the iteration buffers: the footprint is the larger phase
        j->run_st = j->st;
        recon_begin(j.get());
        release_grid_extras(j.get());
        if (!r->ev_ready) GF_HIP(hipEventCreateWithFlags(&r->ev_ready, hipEventDisableTiming));
        if (!r->ev_done) GF_HIP(hipEventCreateWithFlags(&r->ev_done, hipEventDisableTiming));
        GF_HIP(hipEventRecord(r->ev_ready, j->st));
        r->group = std::make_tuple(j->n1, j->n2, j->n1, j->rs.half ? 1 : 0);
        std::lock_guard<std::mutex> lk(mu);
        r->job = std::move(j);
        r->it = 1;
        r->joined = false;
        r->state = GENFIRE_REQ_ITERATING;
        r->t_ready = now_s();
        cv_sched.notify_one();
    } catch (const Error& e) {
        if (e.code == GENFIRE_ERR_NOMEM) {
            // out of device memory after all: back to the queue head until a job finishes
            r->job.reset();
            std::lock_guard<std::mutex> lk(mu);
            used -= r->footprint;
            active.erase(std::remove(active.begin(), active.end(), r), active.end());
            r->state = GENFIRE_REQ_QUEUED;
            queue.push_front(r);
            mem_blocked = !active.empty();
            cv_sched.notify_one();
            return;
        }
        fail(r, e.code, e.msg);
    } catch (const std::exception& e) {
        fail(r, GENFIRE_ERR_ARG, e.what());
    }
}

void genfire_server::finish(genfire_request* r)
{
    try {
        GF_HIP(hipSetDevice(ctx->device));
        genfire_job* j = r->job.get();
        GF_HIP(hipStreamWaitEvent(j->st, r->ev_done, 0));
        j->run_st = j->st;
        recon_end(j);
        r->recon.resize((size_t)box_size(j));
        GF_HIP(hipMemcpy(r->recon.data(), j->recon.p, j->recon.bytes, hipMemcpyDeviceToHost));
        r->errK = j->errK_h;
        r->errR = j->errR_h;
        r->job.reset();   // frees the job's device memory
        std::lock_guard<std::mutex> lk(mu);
        r->state = GENFIRE_REQ_DONE;
        r->t_done = now_s();
        used -= r->footprint;
        active.erase(std::remove(active.begin(), active.end(), r), active.end());
        ++completed;
        lat.push_back(r->t_done - r->t_submit);
        t_last = r->t_done;
        mem_blocked = false;
        cv_done.notify_all();
        cv_sched.notify_one();
    } catch (const Error& e) {
        fail(r, e.code, e.msg);
    } catch (const std::exception& e) {
        fail(r, GENFIRE_ERR_ARG, e.what());
    }
}

// One iteration for every group whose stream has room; true if anything was
// launched.  A group's requests advance (it, FINISHING) only after every launch
// and event record of their iteration succeeded; on any error the group's
// stream is drained and its requests fail with the error, so no request stays
// ITERATING and none is iterated past numIterations.
bool genfire_server::advance_groups()
{
    std::map<std::tuple<int, int, int, int>, std::vector<genfire_request*>> ready;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (genfire_request* r : active)
            if (r->state == GENFIRE_REQ_ITERATING) ready[r->group].push_back(r);
    }
    bool launched = false;
    for (auto& kv : ready) {
        std::vector<genfire_request*>& rs = kv.second;
        Group& g = groups[kv.first];
        int code = GENFIRE_OK;
        std::string msg;
        try {
            if (!g.st) GF_HIP(hipStreamCreateWithFlags(&g.st, hipStreamNonBlocking));
            while (!g.inflight.empty() && hipEventQuery(g.inflight.front()) == hipSuccess) {
                g.pool.push_back(g.inflight.front());
                g.inflight.pop_front();
            }
            if ((int)g.inflight.size() >= cfg.lookahead) continue;
            std::vector<genfire_job*> js;
            std::vector<int> its;
            for (genfire_request* r : rs) {
                genfire_job* j = r->job.get();
                if (!r->joined) {
                    GF_HIP(hipStreamWaitEvent(g.st, r->ev_ready, 0));
                    j->run_st = g.st;
                    r->joined = true;
                }
                js.push_back(j);
                its.push_back(r->it);
            }
            // the group's transform passes (the errK reduction included); for a
            // steady iteration of every member, exactly what the graph holds
            auto body = [&] {
                if (js[0]->rs.half) {
                    for (size_t c = 0; c < js.size(); c += (size_t)cfg.max_batch) {
                        const int n = (int)std::min(js.size() - c, (size_t)cfg.max_batch);
                        half_body_batch(js.data() + c, its.data() + c, n, g.st);
                    }
                } else {
                    for (size_t i = 0; i < js.size(); ++i) {
                        recon_body(js[i], its[i]);
                        recon_post(js[i], its[i]);   // errK only: no R-factor in a steady iteration
                    }
                }
            };
            bool steady = graphs;
            for (size_t i = 0; i < js.size() && steady; ++i) steady = steady_iteration(js[i], its[i]);
            if (steady) {
                std::vector<int64_t> ids;
                for (genfire_request* r : rs) ids.push_back(r->id);
                if (!g.gexec || g.gids != ids) {
                    if (g.gexec) {   // launches of the old graph finish before it goes
                        GF_HIP(hipStreamSynchronize(g.st));
                        g.drop_graph();
                    }
                    GF_HIP(hipStreamBeginCapture(g.st, hipStreamCaptureModeThreadLocal));
                    try {
                        body();
                    } catch (...) {
                        hipGraph_t discard = nullptr;
                        (void)hipStreamEndCapture(g.st, &discard);
                        if (discard) (void)hipGraphDestroy(discard);
                        throw;
                    }
                    GF_HIP(hipStreamEndCapture(g.st, &g.graph));
                    GF_HIP(hipGraphInstantiate(&g.gexec, g.graph, nullptr, nullptr, 0));
                    g.gids = ids;
                }
                GF_HIP(hipGraphLaunch(g.gexec, g.st));
            } else {
                for (size_t i = 0; i < js.size(); ++i) recon_pre(js[i], its[i]);
                body();
                if (js[0]->rs.half)
                    for (size_t i = 0; i < js.size(); ++i) recon_post(js[i], its[i]);   // the R-factor monitor
            }
            if (fail_at > 0 && batched_launches + 1 == fail_at) {
                const int64_t k = fail_at;
                fail_at = 0;   // once
                raise(GENFIRE_ERR_HIP, "server: failure injected at scheduler iteration " + std::to_string(k));
            }
            std::vector<genfire_request*> done;
            for (genfire_request* r : rs)
                if (r->it + 1 > r->job->rs.N) {
                    GF_HIP(hipEventRecord(r->ev_done, g.st));
                    done.push_back(r);
                }
            hipEvent_t ev;
            if (!g.pool.empty()) {
                ev = g.pool.back();
                g.pool.pop_back();
            } else {
                GF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            }
            const hipError_t er = hipEventRecord(ev, g.st);
            if (er != hipSuccess) {
                g.pool.push_back(ev);
                GF_HIP(er);
            }
            g.inflight.push_back(ev);
            launched = true;
            std::lock_guard<std::mutex> lk(mu);
            ++batched_launches;
            batched_jobs += (int64_t)rs.size();
            if (steady) ++graph_replays;
            for (genfire_request* r : rs) ++r->it;
            for (genfire_request* r : done) {
                r->state = GENFIRE_REQ_FINISHING;
                post_locked([this, r] { finish(r); });
            }
        } catch (const Error& e) {
            code = e.code;
            msg = e.msg;
        } catch (const std::bad_alloc&) {
            code = GENFIRE_ERR_NOMEM;
            msg = "host allocation failed";
        } catch (const std::exception& e) {
            code = GENFIRE_ERR_ARG;
            msg = e.what();
        }
        if (code != GENFIRE_OK) {
            // launches already queued for these jobs finish before their memory is freed
            if (g.st) (void)hipStreamSynchronize(g.st);
            g.drop_graph();
            for (genfire_request* r : rs) fail(r, code, msg);
        }
    }
    return launched;
}

void genfire_server::sched_loop()
{
    (void)hipSetDevice(ctx->device);
    for (;;) {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (stop) break;
            admit_locked();
        }
        bool launched = false;
        try {
            launched = advance_groups();
        } catch (const std::exception&) {   // only the grouping itself can throw here (host allocation)
            launched = false;
        }
        std::unique_lock<std::mutex> lk(mu);
        if (stop) break;
        if (!launched) cv_sched.wait_for(lk, std::chrono::microseconds(200));
    }
}

namespace {

genfire_request* find_req(genfire_server* s, int64_t id)
{
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int genfire_server_config_default(genfire_server_config* c)
{
    if (!c) return GENFIRE_ERR_ARG;
    std::memset(c, 0, sizeof(*c));
    c->hbm_budget_bytes = 0;
    c->max_active = 4;
    c->max_batch = kHalfBatchMax;
    c->lookahead = 2;
    c->prep_threads = 2;
    return GENFIRE_OK;
}

int genfire_server_create(genfire_ctx* ctx, const genfire_server_config* cfg, genfire_server** out)
{
    if (!ctx || !out) return GENFIRE_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        use_device(ctx);
        auto s = std::make_unique<genfire_server>();
        s->ctx = ctx;
        if (cfg) s->cfg = *cfg;
        else genfire_server_config_default(&s->cfg);
        genfire_server_config& c = s->cfg;
        GF_CHECK(c.max_active >= 1 && c.max_active <= 64, GENFIRE_ERR_ARG, "server: max_active must be 1..64");
        GF_CHECK(c.max_batch >= 1 && c.max_batch <= kHalfBatchMax, GENFIRE_ERR_ARG, "server: max_batch must be 1..8");
        GF_CHECK(c.lookahead >= 1 && c.lookahead <= 16, GENFIRE_ERR_ARG, "server: lookahead must be 1..16");
        GF_CHECK(c.prep_threads >= 1 && c.prep_threads <= 8, GENFIRE_ERR_ARG, "server: prep_threads must be 1..8");
        if (c.hbm_budget_bytes <= 0) {
            size_t fr = 0, tot = 0;
            GF_HIP(hipMemGetInfo(&fr, &tot));
            c.hbm_budget_bytes = (int64_t)((double)fr * 0.9);
        }
        if (const char* fa = std::getenv("GENFIRE_SERVER_FAIL_AT")) s->fail_at = std::atoll(fa);
        if (const char* ge = std::getenv("GENFIRE_GRAPHS")) s->graphs = std::string(ge) != "0";
        genfire_server* sp = s.get();
        for (int i = 0; i < c.prep_threads + 1; ++i) sp->workers.emplace_back([sp] { sp->worker_loop(); });
        sp->sched = std::thread([sp] { sp->sched_loop(); });
        *out = s.release();
    });
}

int genfire_server_destroy(genfire_server* s)
{
    if (!s) return GENFIRE_OK;
    genfire_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    s->cv_task.notify_all();
    if (s->sched.joinable()) s->sched.join();
    for (auto& t : s->workers)
        if (t.joinable()) t.join();
    (void)hipSetDevice(s->ctx->device);
    for (auto& kv : s->groups) {
        if (kv.second.st) (void)hipStreamSynchronize(kv.second.st);
        for (hipEvent_t e : kv.second.inflight) (void)hipEventDestroy(e);
        for (hipEvent_t e : kv.second.pool) (void)hipEventDestroy(e);
        kv.second.drop_graph();
        if (kv.second.st) (void)hipStreamDestroy(kv.second.st);
    }
    delete s;
    return GENFIRE_OK;
}

int genfire_server_submit(genfire_server* s, const genfire_params* p, int32_t dim1, int32_t dim2, int32_t num_projs,
                          const float* projections, const double* angles, const uint8_t* support,
                          int64_t* request_id)
{
    if (!s || !p || !projections || !angles || !request_id) return GENFIRE_ERR_ARG;
    return guarded(s->ctx, [&] {
        // the checks of genfire_job_create, so that a bad request fails at submission
        validate_job(p, dim1, dim2, num_projs);
        auto r = std::make_unique<genfire_request>();
        r->prm = *p;
        r->dim1 = dim1;
        r->dim2 = dim2;
        r->P = num_projs;
        r->footprint = job_footprint(*p, dim1, dim2, num_projs);
        GF_CHECK(r->footprint <= s->cfg.hbm_budget_bytes, GENFIRE_ERR_NOMEM,
                 "server: request footprint " + std::to_string(r->footprint >> 20) + " MiB exceeds the budget");
        r->proj.assign(projections, projections + (size_t)dim1 * dim2 * num_projs);
        r->ang.assign(angles, angles + 3 * (size_t)num_projs);
        r->has_sup = support != nullptr;
        if (support) r->sup.assign(support, support + (size_t)dim1 * dim2 * dim1);
        r->t_submit = now_s();
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        *request_id = r->id;
        if (s->submitted == 0) s->t_first = r->t_submit;
        ++s->submitted;
        s->queue.push_back(r.get());
        s->reqs[r->id] = std::move(r);
        s->cv_sched.notify_one();
    });
}

int genfire_server_status(genfire_server* s, int64_t id, int32_t* state)
{
    if (!s || !state) return GENFIRE_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    genfire_request* r = find_req(s, id);
    if (!r) return GENFIRE_ERR_ARG;
    *state = r->state;
    return GENFIRE_OK;
}

int genfire_server_wait(genfire_server* s, int64_t id)
{
    if (!s) return GENFIRE_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    genfire_request* r = find_req(s, id);
    if (!r) return GENFIRE_ERR_ARG;
    s->cv_done.wait(lk, [&] { return r->state == GENFIRE_REQ_DONE || r->state == GENFIRE_REQ_FAILED; });
    if (r->state == GENFIRE_REQ_FAILED) {
        set_err(s->ctx, r->err);
        return r->status;
    }
    return GENFIRE_OK;
}

int genfire_server_drain(genfire_server* s)
{
    if (!s) return GENFIRE_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->active.empty(); });
    return GENFIRE_OK;
}

int genfire_server_result(genfire_server* s, int64_t id, float* reconstruction, float* errK, float* errR,
                          double latency_ms[4])
{
    if (!s) return GENFIRE_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    genfire_request* r = find_req(s, id);
    if (!r) return GENFIRE_ERR_ARG;
    if (r->state == GENFIRE_REQ_FAILED) {
        set_err(s->ctx, r->err);
        return r->status;
    }
    if (r->state != GENFIRE_REQ_DONE) return GENFIRE_ERR_STATE;
    if (reconstruction) std::memcpy(reconstruction, r->recon.data(), sizeof(float) * r->recon.size());
    if (errK) std::memcpy(errK, r->errK.data(), sizeof(float) * r->errK.size());
    if (errR) std::memcpy(errR, r->errR.data(), sizeof(float) * r->errR.size());
    if (latency_ms) {
        latency_ms[0] = (r->t_admit - r->t_submit) * 1e3;   // queued
        latency_ms[1] = (r->t_ready - r->t_admit) * 1e3;    // preparing (upload, gridding, set-up)
        latency_ms[2] = (r->t_done - r->t_ready) * 1e3;     // iterating and read-back
        latency_ms[3] = (r->t_done - r->t_submit) * 1e3;    // time to reconstruction
    }
    return GENFIRE_OK;
}

int genfire_server_release(genfire_server* s, int64_t id)
{
    if (!s) return GENFIRE_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    genfire_request* r = find_req(s, id);
    if (!r) return GENFIRE_ERR_ARG;
    if (r->state != GENFIRE_REQ_DONE && r->state != GENFIRE_REQ_FAILED) return GENFIRE_ERR_STATE;
    s->reqs.erase(id);
    return GENFIRE_OK;
}

int genfire_server_get_stats(genfire_server* s, genfire_server_stats* out)
{
    if (!s || !out) return GENFIRE_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->queued = (int64_t)s->queue.size();
    out->active = (int64_t)s->active.size();
    out->max_active_seen = s->max_active_seen;
    out->batched_launches = s->batched_launches;
    out->batched_jobs = s->batched_jobs;
    out->graph_replays = s->graph_replays;
    out->peak_footprint_bytes = s->peak_used;
    out->hbm_budget_bytes = s->cfg.hbm_budget_bytes;
    if (!s->lat.empty()) {
        std::vector<double> v = s->lat;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            // linear interpolation between closest ranks (numpy's default)
            const double pos = q * (double)(v.size() - 1);
            const size_t lo = (size_t)pos;
            const size_t hi = std::min(v.size() - 1, lo + 1);
            return (v[lo] + (pos - (double)lo) * (v[hi] - v[lo])) * 1e3;
        };
        out->p50_ms = pct(0.50);
        out->p99_ms = pct(0.99);
        double sum = 0.0;
        for (double x : v) sum += x;
        out->mean_ms = sum / (double)v.size() * 1e3;
        const double span = s->t_last - s->t_first;
        out->jobs_per_hour = span > 0 ? (double)s->completed / span * 3600.0 : 0.0;
    }
    return GENFIRE_OK;
}

}  // extern "C"
