// Batches (many prepared meshes iterated by the same launches, one launch per kernel per
// level covering every mesh, replayed from one captured hipGraph) and the drag-request server.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "md_runtime.h"

using namespace md;

struct md_batch {
    md_ctx *ctx = nullptr;
    std::vector<md_mesh *> meshes;
    std::vector<uint64_t> gens;  // device layout generation of each mesh when the plan was built
    Plan plan;
    bool use_graph = true;
};

namespace {

hipError_t batch_plan(md_batch *b) {
    std::vector<const MeshDev *> devs;
    std::vector<const Factor *> facs;
    b->gens.clear();
    for (md_mesh *m : b->meshes) {
        devs.push_back(&m->dev);
        facs.push_back(&m->fac);
        b->gens.push_back(m->gen);
    }
    const md_config &c = b->meshes[0]->cfg;
    return build_plan(b->plan, devs, facs, c.precision == MD_PREC_FP64, c.refine_steps, b->ctx->stream, true, true);
}

// A mesh re-prepared with a new constraint set, or given another sweep layout, since the
// plan was built has moved its device arrays: launching the old plan would read freed memory.
int batch_current(md_batch *b, const char *who) {
    for (size_t k = 0; k < b->meshes.size(); ++k)
        if (b->meshes[k]->gen != b->gens[k] || b->meshes[k]->dirty) {
            set_error(std::string(who) + ": mesh " + std::to_string(k) +
                      " was re-prepared or re-laid out after the batch was built");
            return MD_ERR_STATE;
        }
    return MD_OK;
}

}  // namespace

extern "C" int md_batch_create(md_ctx *ctx, int count, md_mesh *const *meshes, md_batch **out) {
    if (!ctx || !out || count <= 0 || !meshes) { set_error("md_batch_create: invalid argument"); return MD_ERR_INVALID; }
    *out = nullptr;
    std::unique_ptr<md_batch> b(new md_batch());
    b->ctx = ctx;
    int prec = meshes[0]->cfg.precision, refine = meshes[0]->cfg.refine_steps;
    for (int k = 0; k < count; ++k) {
        md_mesh *m = meshes[k];
        if (!m || m->ctx != ctx) { set_error("md_batch_create: mesh from another context"); return MD_ERR_INVALID; }
        if (m->dirty || !m->analyzed) { set_error("md_batch_create: mesh " + std::to_string(k) + " is not prepared"); return MD_ERR_STATE; }
        if (m->cfg.precision != prec || m->cfg.refine_steps != refine) {
            set_error("md_batch_create: meshes differ in precision or refinement steps");
            return MD_ERR_INVALID;
        }
        b->meshes.push_back(m);
    }
    b->use_graph = meshes[0]->cfg.use_graph != 0;
    for (int k = 0; k < count; ++k)
        if (meshes[k]->cfg.solver != meshes[0]->cfg.solver || meshes[k]->cfg.pcg_max_iters != meshes[0]->cfg.pcg_max_iters) {
            set_error("md_batch_create: meshes differ in solver settings");
            return MD_ERR_INVALID;
        }
    b->plan.pcg = meshes[0]->cfg.solver == MD_SOLVER_PCG;
    b->plan.pcg_iters = meshes[0]->cfg.pcg_max_iters;
    hipSetDevice(ctx->device);
    hipError_t e = batch_plan(b.get());
    if (e != hipSuccess) { set_error(std::string("md_batch_create: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    ctx_retain(ctx);
    *out = b.release();
    return MD_OK;
}

extern "C" int md_batch_iterate(md_batch *b, int iterations) {
    if (!b || iterations < 0) { set_error("md_batch_iterate: invalid argument"); return MD_ERR_INVALID; }
    if (int rc = batch_current(b, "md_batch_iterate")) return rc;
    hipSetDevice(b->ctx->device);
    hipError_t e = run_iterations(b->plan, b->ctx->stream, iterations, b->use_graph);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) { set_error(std::string("md_batch_iterate: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    for (md_mesh *m : b->meshes) m->iterated = true;
    return MD_OK;
}

extern "C" int md_batch_time_iterations(md_batch *b, int warmup, int iterations, double *total_ms) {
    if (!b || iterations < 0 || warmup < 0) { set_error("md_batch_time_iterations: invalid argument"); return MD_ERR_INVALID; }
    if (int rc = batch_current(b, "md_batch_time_iterations")) return rc;
    hipSetDevice(b->ctx->device);
    hipStream_t st = b->ctx->stream;
    hipError_t e = run_iterations(b->plan, st, warmup, b->use_graph);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipEvent_t a = nullptr, c = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&c);
    if (e == hipSuccess) e = hipEventRecord(a, st);
    if (e == hipSuccess) e = run_iterations(b->plan, st, iterations, b->use_graph);
    if (e == hipSuccess) e = hipEventRecord(c, st);
    if (e == hipSuccess) e = hipEventSynchronize(c);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, c);
    if (a) hipEventDestroy(a);
    if (c) hipEventDestroy(c);
    if (e != hipSuccess) { set_error(std::string("md_batch_time_iterations: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    if (total_ms) *total_ms = ms;
    return MD_OK;
}

extern "C" int md_batch_autotune(md_batch *b, const int *budgets, int count, int iterations, int *best_budget,
                                 double *best_ms) {
    if (!b) { set_error("md_batch_autotune: null batch"); return MD_ERR_INVALID; }
    if (!budgets || count <= 0 || iterations <= 0) {
        set_error("md_batch_autotune: need at least one candidate and one iteration");
        return MD_ERR_INVALID;
    }
    for (int i = 0; i < count; ++i)
        if (budgets[i] < 0 || budgets[i] > kSubBudgetMax) {
            set_error("md_batch_autotune: budget " + std::to_string(budgets[i]) + " outside [0, " +
                      std::to_string(kSubBudgetMax) + "]");
            return MD_ERR_INVALID;
        }
    if (int rc = batch_current(b, "md_batch_autotune")) return rc;
    if (b->meshes[0]->cfg.solver != MD_SOLVER_LDLT) {
        set_error("md_batch_autotune: only the LDLT solver has a sweep layout");
        return MD_ERR_INVALID;
    }
    for (md_mesh *m : b->meshes)
        if (m->servers > 0) { set_error("md_batch_autotune: a mesh is registered with a server"); return MD_ERR_STATE; }
    hipSetDevice(b->ctx->device);
    hipStream_t st = b->ctx->stream;
    StateSnapshot snap;
    hipError_t e = snapshot_state(b->meshes, st, snap);
    if (e != hipSuccess) { set_error(std::string("md_batch_autotune: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    int best = b->meshes[0]->cfg.subtree_budget;
    double best_t = 1e300;
    auto use = [&](int budget) -> int {
        for (md_mesh *m : b->meshes)
            if (int rc = relayout(m, budget)) return rc;
        hipError_t x = batch_plan(b);
        if (x != hipSuccess) { set_error(std::string("md_batch_autotune: ") + hipGetErrorString(x)); return MD_ERR_HIP; }
        return MD_OK;
    };
    for (int i = 0; i < count; ++i) {
        if (int rc = use(budgets[i])) return rc;
        double t = 0.0;
        e = run_iterations(b->plan, st, 1, b->use_graph);  // first replay captures the graph
        if (e == hipSuccess) e = restore_state(b->meshes, st, snap);
        if (e == hipSuccess) e = time_plan(b->plan, st, iterations, b->use_graph, t);
        if (e == hipSuccess) e = restore_state(b->meshes, st, snap);
        if (e != hipSuccess) { set_error(std::string("md_batch_autotune: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
        if (t < best_t) { best_t = t; best = budgets[i]; }
    }
    if (int rc = use(best)) return rc;
    if (best_budget) *best_budget = best;
    if (best_ms) *best_ms = best_t / iterations;
    return MD_OK;
}

extern "C" int md_batch_destroy(md_batch *b) {
    if (!b) return MD_OK;
    md_ctx *ctx = b->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    delete b;
    ctx_release(ctx);
    return MD_OK;
}

// ---- server (continuous batching) ---------------------------------------------------------
//
// One worker thread owns the context's stream while the server runs. Requests queue in
// submission order; at every iteration boundary the worker admits queued requests whose mesh
// is idle (one request in flight per mesh, at most max_active in flight), rebuilds the launch
// plan when the set changed, and runs one ARAP iteration for every admitted request in the
// same launches. A request leaves the set when its iterations are done; its positions are
// downloaded and its waiters woken. Meshes are resident from md_server_add_mesh on, within the
// HBM budget.

namespace {
using SClock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    int mesh_id = -1;
    std::vector<int> idx;
    std::vector<double> xyz;
    int iters = 0, left = 0;
    SClock::time_point t_submit, t_done;
    bool admitted = false, done = false;
    int status = MD_OK;
    std::string error;
    std::vector<int> iv;        // internal ids of the moved targets (upload source)
    d4 *pinned = nullptr;       // positions in internal order, copied at completion
    int n = 0;
};
}  // namespace

struct md_server {
    md_ctx *ctx = nullptr;
    md_server_config cfg{};
    std::vector<md_mesh *> meshes;
    size_t resident = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<uint64_t, std::shared_ptr<Request>> tickets;
    std::vector<std::shared_ptr<Request>> active;
    std::vector<char> busy;  // per mesh
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    Plan plan;
    std::vector<int> plan_meshes;
    std::string fatal;
    std::vector<std::vector<d4 *>> pinned_free;  // per mesh: reusable pinned result buffers
    std::vector<d4 *> pinned_all;

    d4 *take_pinned(int mesh_id, int n) {  // caller holds mu
        auto &fl = pinned_free[mesh_id];
        if (!fl.empty()) {
            d4 *p = fl.back();
            fl.pop_back();
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, sizeof(d4) * (size_t)std::max(1, n), hipHostMallocDefault) != hipSuccess) return nullptr;
        pinned_all.push_back((d4 *)p);
        return (d4 *)p;
    }
};

namespace {

// Pastes the request's targets into p' (fixed rest pose: the previous result is the initial
// guess): one upload and one scatter launch, stream-ordered, no host synchronisation. The
// sources live in the request until it completes.
int apply_targets(md_mesh *m, Request &r) {
    int k = (int)r.idx.size();
    if (k == 0) return MD_OK;
    r.iv.resize(k);
    for (int i = 0; i < k; ++i) {
        r.iv[i] = m->topo.internal_of[r.idx[i]];
        std::memcpy(&m->con[3 * (size_t)r.idx[i]], &r.xyz[3 * (size_t)i], 3 * sizeof(double));
    }
    hipStream_t st = m->ctx->stream;
    hipError_t e = hipMemcpyAsync(m->d_tgt_idx, r.iv.data(), sizeof(int) * k, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->d_tgt_xyz, r.xyz.data(), sizeof(double) * 3 * k, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        launch_scatter_targets(st, m->dev.pp, m->d_tgt_idx, m->d_tgt_xyz, k);
        e = hipGetLastError();
    }
    return e == hipSuccess ? MD_OK : MD_ERR_HIP;
}

void server_loop(md_server *S) {
    hipSetDevice(S->ctx->device);
    hipStream_t st = S->ctx->stream;
    for (;;) {
        std::vector<std::shared_ptr<Request>> admitted_now;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) return;
                // admission at the iteration boundary
                for (auto it = S->queue.begin(); it != S->queue.end();) {
                    auto &r = *it;
                    bool slot = S->cfg.max_active <= 0 || (int)S->active.size() < S->cfg.max_active;
                    if (slot && !S->busy[r->mesh_id]) {
                        S->busy[r->mesh_id] = 1;
                        r->admitted = true;
                        S->active.push_back(r);
                        admitted_now.push_back(r);
                        it = S->queue.erase(it);
                    } else {
                        ++it;
                    }
                }
                if (!S->active.empty()) break;
                S->cv_work.wait(lk);
            }
            // pinned result buffers for the requests that finish after this iteration
            for (auto &r : S->active)
                if (r->left <= 1 && !r->pinned && r->status == MD_OK) {
                    r->n = S->meshes[r->mesh_id]->n;
                    r->pinned = S->take_pinned(r->mesh_id, r->n);
                    if (!r->pinned) { r->status = MD_ERR_NOMEM; r->error = "pinned result buffer"; r->left = 0; }
                }
        }
        for (auto &r : admitted_now) {
            int rc = apply_targets(S->meshes[r->mesh_id], *r);
            if (rc) { r->status = rc; r->error = "target upload failed"; r->left = 0; }
        }
        // plan over the active meshes (rebuilt only when the set changed; stream-ordered)
        std::vector<int> ids;
        for (auto &r : S->active) ids.push_back(r->mesh_id);
        if (ids != S->plan_meshes) {
            std::vector<const MeshDev *> devs;
            std::vector<const Factor *> facs;
            for (int id : ids) { devs.push_back(&S->meshes[id]->dev); facs.push_back(&S->meshes[id]->fac); }
            md_mesh *m0 = S->meshes[ids[0]];
            S->plan.pcg = m0->cfg.solver == MD_SOLVER_PCG;
            S->plan.pcg_iters = m0->cfg.pcg_max_iters;
            hipError_t e = build_plan(S->plan, devs, facs, m0->fp64(), m0->cfg.refine_steps, st, false);
            if (e != hipSuccess) { S->fatal = hipGetErrorString(e); }
            S->plan_meshes = ids;
        }
        bool any_left = false;
        for (auto &r : S->active) any_left |= r->left > 0;
        hipError_t e = hipSuccess;
        if (any_left && S->fatal.empty()) {
            enqueue_iteration(S->plan, st, nullptr);
            e = hipGetLastError();
        }
        // results of the requests finishing now: device-to-host into their pinned buffers
        for (auto &r : S->active)
            if (e == hipSuccess && r->left <= 1 && r->pinned && r->status == MD_OK) {
                md_mesh *m = S->meshes[r->mesh_id];
                e = hipMemcpyAsync(r->pinned, m->dev.pp, sizeof(d4) * m->n, hipMemcpyDeviceToHost, st);
            }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        std::vector<std::shared_ptr<Request>> finished;
        for (auto &r : S->active) {
            if (r->left > 0) r->left--;
            if (e != hipSuccess || !S->fatal.empty()) {
                r->status = MD_ERR_HIP;
                r->error = e != hipSuccess ? hipGetErrorString(e) : S->fatal;
                r->left = 0;
            }
            if (r->left == 0) finished.push_back(r);
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            auto now = SClock::now();
            for (auto &r : finished) {
                S->meshes[r->mesh_id]->iterated = true;
                r->done = true;
                r->t_done = now;
                S->busy[r->mesh_id] = 0;
            }
            S->active.erase(std::remove_if(S->active.begin(), S->active.end(),
                                           [](const std::shared_ptr<Request> &r) { return r->done; }),
                            S->active.end());
        }
        if (!finished.empty()) S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" int md_server_create(md_ctx *ctx, const md_server_config *cfg, md_server **out) {
    if (!ctx || !out) { set_error("md_server_create: invalid argument"); return MD_ERR_INVALID; }
    std::unique_ptr<md_server> S(new md_server());
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    ctx_retain(ctx);
    S->worker = std::thread(server_loop, S.get());
    *out = S.release();
    return MD_OK;
}

extern "C" int md_server_add_mesh(md_server *S, md_mesh *m, int *mesh_id) {
    if (!S || !m || !mesh_id) { set_error("md_server_add_mesh: invalid argument"); return MD_ERR_INVALID; }
    if (m->ctx != S->ctx) { set_error("md_server_add_mesh: mesh from another context"); return MD_ERR_INVALID; }
    if (m->dirty || !m->analyzed || m->cfg.rest_mode != MD_REST_FIXED) {
        set_error("md_server_add_mesh: the mesh must be prepared (md_mesh_prepare) with MD_REST_FIXED");
        return MD_ERR_STATE;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (!S->meshes.empty()) {
        md_mesh *m0 = S->meshes[0];
        if (m0->cfg.precision != m->cfg.precision || m0->cfg.refine_steps != m->cfg.refine_steps ||
            m0->cfg.solver != m->cfg.solver || m0->cfg.pcg_max_iters != m->cfg.pcg_max_iters) {
            set_error("md_server_add_mesh: meshes of one server share precision and refinement steps");
            return MD_ERR_INVALID;
        }
    }
    size_t bytes = m->arena.size() + m->larena.size();
    if (S->cfg.hbm_budget_bytes > 0 && S->resident + bytes > S->cfg.hbm_budget_bytes) {
        set_error("md_server_add_mesh: mesh needs " + std::to_string(bytes) + " device bytes, " +
                  std::to_string(S->cfg.hbm_budget_bytes - S->resident) + " left in the HBM budget");
        return MD_ERR_BUDGET;
    }
    S->resident += bytes;
    S->meshes.push_back(m);
    ++m->servers;
    S->busy.push_back(0);
    S->pinned_free.emplace_back();
    *mesh_id = (int)S->meshes.size() - 1;
    return MD_OK;
}

extern "C" int md_server_submit(md_server *S, int mesh_id, int k, const int *idx, const double *xyz,
                                int iterations, uint64_t *ticket) {
    if (!S || !ticket || k < 0 || (k > 0 && (!idx || !xyz)) || iterations < 0) {
        set_error("md_server_submit: invalid argument");
        return MD_ERR_INVALID;
    }
    auto r = std::make_shared<Request>();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (mesh_id < 0 || mesh_id >= (int)S->meshes.size()) { set_error("md_server_submit: unknown mesh"); return MD_ERR_INVALID; }
        md_mesh *m = S->meshes[mesh_id];
        if (k > m->n - m->topo.nf) { set_error("md_server_submit: more targets than constrained vertices"); return MD_ERR_INVALID; }
        for (int i = 0; i < k; ++i)
            if (idx[i] < 0 || idx[i] >= m->n || !m->is_con[idx[i]]) {
                set_error("md_server_submit: vertex " + std::to_string(idx[i]) + " is not a constrained vertex of the mesh");
                return MD_ERR_INVALID;
            }
        r->ticket = S->next_ticket++;
        r->mesh_id = mesh_id;
        r->idx.assign(idx, idx + k);
        r->xyz.assign(xyz, xyz + 3 * (size_t)k);
        r->iters = r->left = iterations;
        r->t_submit = SClock::now();
        S->queue.push_back(r);
        S->tickets[r->ticket] = r;
        *ticket = r->ticket;
    }
    S->cv_work.notify_one();
    return MD_OK;
}

// Called with S->mu held: takes the finished request out of the ticket map. The permutation
// of its positions into the original vertex order runs afterwards without the lock.
static std::shared_ptr<Request> take_finished(md_server *S, uint64_t ticket) {
    auto it = S->tickets.find(ticket);
    std::shared_ptr<Request> r = it->second;
    S->tickets.erase(it);
    return r;
}

static int deliver(md_server *S, std::shared_ptr<Request> r, double *xyz, double *latency_ms) {
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    int rc = r->status;
    if (rc != MD_OK) set_error("request failed: " + r->error);
    else if (xyz && r->pinned) {
        md_mesh *m = S->meshes[r->mesh_id];
        for (int iv = 0; iv < m->n; ++iv) {
            int v = m->topo.original_of[iv];
            xyz[3 * (size_t)v] = r->pinned[iv].x;
            xyz[3 * (size_t)v + 1] = r->pinned[iv].y;
            xyz[3 * (size_t)v + 2] = r->pinned[iv].z;
        }
    }
    if (r->pinned) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->pinned_free[r->mesh_id].push_back(r->pinned);
        r->pinned = nullptr;
    }
    return rc;
}

extern "C" int md_server_poll(md_server *S, uint64_t ticket, int *done, double *xyz, double *latency_ms) {
    if (!S || !done) { set_error("md_server_poll: invalid argument"); return MD_ERR_INVALID; }
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->tickets.find(ticket);
        if (it == S->tickets.end()) { set_error("md_server_poll: unknown ticket"); return MD_ERR_INVALID; }
        *done = it->second->done ? 1 : 0;
        if (!*done) return MD_OK;
        r = take_finished(S, ticket);
    }
    return deliver(S, r, xyz, latency_ms);
}

extern "C" int md_server_wait(md_server *S, uint64_t ticket, double timeout_ms, double *xyz, double *latency_ms) {
    if (!S) { set_error("md_server_wait: invalid argument"); return MD_ERR_INVALID; }
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->tickets.find(ticket);
        if (it == S->tickets.end()) { set_error("md_server_wait: unknown ticket"); return MD_ERR_INVALID; }
        r = it->second;
        bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), [&] { return r->done; });
        if (!ok) { set_error("md_server_wait: timed out"); return MD_ERR_TIMEOUT; }
        take_finished(S, ticket);
    }
    return deliver(S, r, xyz, latency_ms);
}

extern "C" int md_server_destroy(md_server *S) {
    if (!S) return MD_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    md_ctx *ctx = S->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (d4 *p : S->pinned_all) hipHostFree(p);
    for (md_mesh *m : S->meshes) --m->servers;
    delete S;
    ctx_release(ctx);
    return MD_OK;
}
