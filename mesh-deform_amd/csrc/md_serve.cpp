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

#include "md_admission.h"
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
// A mesh registered with a server belongs to the server's stream until the server is gone.
int batch_current(md_batch *b, const char *who) {
    for (size_t k = 0; k < b->meshes.size(); ++k) {
        if (b->meshes[k]->servers > 0) {
            set_error(std::string(who) + ": mesh " + std::to_string(k) + " is registered with a server");
            return MD_ERR_STATE;
        }
        if (b->meshes[k]->gen != b->gens[k] || b->meshes[k]->dirty || b->meshes[k]->evicted) {
            set_error(std::string(who) + ": mesh " + std::to_string(k) +
                      " was re-prepared, re-laid out or evicted after the batch was built");
            return MD_ERR_STATE;
        }
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
        hipSetDevice(ctx->device);
        if (int rc = ensure_resident(m, "md_batch_create")) return rc;
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
// One worker thread owns the server's stream. Requests queue in submission order; at every
// iteration boundary the worker admits queued requests whose mesh is idle (one request in
// flight per mesh, at most max_active in flight) and resident. A request whose mesh was
// evicted is admitted when its footprint fits the HBM budget, after the least recently used
// idle meshes are evicted to host memory if that makes room; it waits in the queue otherwise,
// until iterations finish and memory frees up. Every admitted request is iterated in the same
// launches (one ARAP iteration per boundary for the whole set); a request leaves the set when
// its iterations are done and its positions are copied to a pinned buffer. Launches run one
// iteration ahead of the host: the host admits, plans and launches iteration i+1 before it waits
// for iteration i, so the device does not idle at a boundary.

namespace {
using SClock = std::chrono::steady_clock;

struct ServedMesh {
    md_mesh *m = nullptr;
    int id = 0;
    size_t bytes = 0;             // device footprint while resident (both arenas)
    bool busy = false;            // a request admitted and not completed
    uint64_t last_use = 0;        // admission stamp (least recently used eviction)
    std::vector<d4 *> pinned_free;  // reusable pinned result buffers
    std::vector<uint32_t> stamp;    // per vertex: the last submission that moved it (dedup)
    uint32_t stamp_gen = 0;
};

struct Request {
    uint64_t ticket = 0;
    ServedMesh *sm = nullptr;     // fixed at submission (the server's mesh table may grow)
    std::vector<int> idx;
    std::vector<double> xyz;
    int iters = 0, left = 0;
    SClock::time_point t_submit, t_done;
    bool done = false;
    int status = MD_OK;
    std::string error;
    d4 *pinned = nullptr;         // positions in internal order, copied at completion
    uint64_t finish_seq = 0;      // launch after which the copy of its positions is complete
};
}  // namespace

struct md_server {
    md_ctx *ctx = nullptr;
    md_server_config cfg{};
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // result copies (default priority, like every stream here)
    std::vector<std::unique_ptr<ServedMesh>> meshes;  // guarded by mu; the entries never move
    size_t resident = 0, peak_resident = 0;
    uint64_t lru_clock = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::map<uint64_t, std::shared_ptr<Request>> tickets;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    md_server_stats stats{};      // guarded by mu
    // worker thread only
    std::vector<std::shared_ptr<Request>> active;     // iterated at the next boundary
    std::vector<std::shared_ptr<Request>> finishing;  // positions being copied
    Plan plans[2];
    std::vector<std::pair<int, uint64_t>> keys[2];    // (mesh id, generation) of each plan's set
    hipEvent_t ev[2] = {nullptr, nullptr};            // after launch seq (by parity)
    hipEvent_t evc[2] = {nullptr, nullptr};           // after the result copies of launch seq
    hipEvent_t ev_flush = nullptr;                    // copies at a boundary without a launch
    uint64_t seq = 0;                                 // launches so far
    uint64_t boundaries = 0;                          // iteration boundaries so far
    // targets of the requests admitted at a boundary: pinned staging by boundary parity (the
    // copy from the buffer of boundary b has completed when boundary b + 2 writes it), one
    // device buffer (stream-ordered)
    TargetEntry *tgt_host[2] = {nullptr, nullptr};
    size_t tgt_cap_host[2] = {0, 0};
    TargetEntry *tgt_dev = nullptr;
    size_t tgt_cap_dev = 0;
    std::vector<d4 *> pinned_all;                     // guarded by mu

    d4 *take_pinned(ServedMesh *sm) {  // caller holds mu
        auto &fl = sm->pinned_free;
        if (!fl.empty()) {
            d4 *p = fl.back();
            fl.pop_back();
            return p;
        }
        void *p = nullptr;
        if (hipHostMalloc(&p, sizeof(d4) * (size_t)std::max(1, sm->m->n), hipHostMallocDefault) != hipSuccess)
            return nullptr;
        pinned_all.push_back((d4 *)p);
        return (d4 *)p;
    }
};

namespace {

// Pastes the targets of the requests admitted at this boundary into their meshes' p' (fixed
// rest pose: the previous result is the initial guess): the entries (address of p'_i, target)
// of every request gathered into this boundary's pinned staging buffer, one upload and one
// scatter launch, stream-ordered, no host synchronisation.
hipError_t apply_targets(md_server *S, const std::vector<std::shared_ptr<Request>> &reqs, hipStream_t st,
                         std::string &what) {
    size_t k = 0;
    for (auto &r : reqs) k += r->idx.size();
    if (k == 0) return hipSuccess;
    const int par = (int)(S->boundaries & 1);
    if (S->tgt_cap_host[par] < k) {
        what = "pinned target staging";
        if (S->tgt_host[par]) hipHostFree(S->tgt_host[par]);
        S->tgt_host[par] = nullptr;
        S->tgt_cap_host[par] = 0;
        const size_t cap = k + k / 2;
        if (hipError_t e = hipHostMalloc((void **)&S->tgt_host[par], sizeof(TargetEntry) * cap, hipHostMallocDefault))
            return e;
        S->tgt_cap_host[par] = cap;
    }
    if (S->tgt_cap_dev < k) {
        // the device buffer is read by earlier scatter launches on this stream only
        what = "device target staging";
        if (hipError_t e = hipStreamSynchronize(st)) return e;
        if (S->tgt_dev) hipFree(S->tgt_dev);
        S->tgt_dev = nullptr;
        S->tgt_cap_dev = 0;
        const size_t cap = k + k / 2;
        if (hipError_t e = hipMalloc((void **)&S->tgt_dev, sizeof(TargetEntry) * cap)) return e;
        S->tgt_cap_dev = cap;
    }
    TargetEntry *h = S->tgt_host[par];
    size_t n = 0;
    for (auto &r : reqs) {
        md_mesh *m = r->sm->m;
        for (size_t i = 0; i < r->idx.size(); ++i, ++n) {
            const int v = r->idx[i];
            const double *x = &r->xyz[3 * i];
            std::memcpy(&m->con[3 * (size_t)v], x, 3 * sizeof(double));
            h[n].addr = reinterpret_cast<unsigned long long>(m->dev.pp + m->topo.internal_of[v]);
            h[n].x = x[0];
            h[n].y = x[1];
            h[n].z = x[2];
        }
    }
    what = "target upload";
    hipError_t e = hipMemcpyAsync(S->tgt_dev, h, sizeof(TargetEntry) * k, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        what = "target scatter launch";
        launch_scatter_entries(st, S->tgt_dev, (int)k);
        e = hipGetLastError();
    }
    return e;
}

void fail(Request &r, int status, const std::string &msg) {
    if (r.status == MD_OK) {
        r.status = status;
        r.error = msg;
    }
    r.left = 0;
}

// Admission at an iteration boundary (S->mu held), by plan_admission over the served meshes'
// state. Fills the evictions and restorations it needs; admitted requests join S->active with
// their meshes marked busy.
void admit(md_server *S, std::vector<ServedMesh *> &evict, std::vector<ServedMesh *> &restore,
           std::vector<std::shared_ptr<Request>> &admitted) {
    if (S->queue.empty()) return;
    std::vector<AdmitMesh> ms(S->meshes.size());
    for (size_t k = 0; k < ms.size(); ++k) {
        const ServedMesh &x = *S->meshes[k];
        ms[k].bytes = x.bytes;
        ms[k].last_use = x.last_use;
        ms[k].resident = !x.m->evicted;
        ms[k].busy = x.busy;
    }
    std::vector<int> queue;
    queue.reserve(S->queue.size());
    for (auto &r : S->queue) queue.push_back(r->sm->id);
    size_t resident = S->resident;
    std::vector<char> take;
    std::vector<int> ev, rs;
    plan_admission(ms, queue, (int)S->active.size(), S->cfg.max_active, S->cfg.hbm_budget_bytes, resident,
                   S->lru_clock, take, ev, rs);
    for (int k : ev) evict.push_back(S->meshes[k].get());
    for (int k : rs) restore.push_back(S->meshes[k].get());
    std::deque<std::shared_ptr<Request>> left;
    for (size_t q = 0; q < take.size(); ++q) {
        std::shared_ptr<Request> &r = S->queue[q];
        if (!take[q]) {
            left.push_back(r);
            continue;
        }
        r->sm->busy = true;
        r->sm->last_use = ms[r->sm->id].last_use;
        S->active.push_back(r);
        admitted.push_back(r);
    }
    S->queue.swap(left);
}

// Marks requests done (S->mu not held) and wakes their waiters.
void complete(md_server *S, std::vector<std::shared_ptr<Request>> &done) {
    if (done.empty()) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto now = SClock::now();
        for (auto &r : done) {
            r->sm->m->iterated = true;
            r->done = true;
            r->t_done = now;
            r->sm->busy = false;
            if (r->status == MD_OK) ++S->stats.completed;
            else ++S->stats.failed;
        }
    }
    done.clear();
    S->cv_done.notify_all();
}

// Waits for the launches up to `upto` and the copies of their results (all queued work of both
// streams when upto is 0) and completes the requests whose positions have arrived.
void drain(md_server *S, uint64_t upto, hipError_t &err) {
    hipError_t e = hipSuccess;
    if (upto) {
        e = hipEventSynchronize(S->ev[upto % 2]);
        if (e == hipSuccess) e = hipEventSynchronize(S->evc[upto % 2]);
    } else {
        e = hipStreamSynchronize(S->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(S->copy_stream);
    }
    if (e != hipSuccess && err == hipSuccess) err = e;
    std::vector<std::shared_ptr<Request>> done;
    for (auto &r : S->finishing)
        if (upto == 0 || r->finish_seq <= upto) {
            if (err != hipSuccess) fail(*r, MD_ERR_HIP, hipGetErrorString(err));
            done.push_back(r);
        }
    S->finishing.erase(std::remove_if(S->finishing.begin(), S->finishing.end(),
                                      [&](const std::shared_ptr<Request> &r) {
                                          return std::find(done.begin(), done.end(), r) != done.end();
                                      }),
                       S->finishing.end());
    complete(S, done);
}

void server_loop(md_server *S) {
    hipSetDevice(S->ctx->device);
    hipStream_t st = S->stream;
    int cur = 0;  // plan of the last launch
    for (;;) {
        std::vector<ServedMesh *> evict, restore;
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            for (;;) {
                if (S->stop) break;
                admit(S, evict, restore, admitted);
                if (!S->active.empty() || !S->finishing.empty()) break;
                S->cv_work.wait(lk);
            }
            if (S->stop) {
                lk.unlock();
                hipError_t e = hipSuccess;
                drain(S, 0, e);
                std::vector<std::shared_ptr<Request>> left;
                {
                    std::lock_guard<std::mutex> g(S->mu);
                    for (auto &r : S->queue) left.push_back(r);
                    S->queue.clear();
                }
                for (auto &r : S->active) left.push_back(r);
                S->active.clear();
                for (auto &r : left) fail(*r, MD_ERR_STATE, "the server was destroyed");
                complete(S, left);
                return;
            }
        }
        // residency changes: the victims are idle, so no launch in flight reads them
        for (ServedMesh *v : evict) {
            int rc = mesh_evict(v->m, st);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc == MD_OK) {
                S->resident -= v->bytes;
                ++S->stats.evictions;
            }
        }
        for (ServedMesh *sm : restore) {
            int rc = mesh_restore(sm->m, st);
            std::string msg = rc ? md_last_error() : "";
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc == MD_OK) {
                S->resident += sm->bytes;
                S->peak_resident = std::max(S->peak_resident, S->resident);
                ++S->stats.restores;
            } else {
                for (auto &r : admitted)
                    if (r->sm == sm) fail(*r, rc, "restoring the mesh: " + msg);
            }
        }
        {
            std::vector<std::shared_ptr<Request>> moving;
            for (auto &r : admitted) {
                if (r->status != MD_OK) continue;
                if (r->sm->m->evicted) {  // its restoration did not happen: nothing may launch on it
                    fail(*r, MD_ERR_STATE, "the mesh is not resident");
                    continue;
                }
                moving.push_back(r);
            }
            std::string what;
            hipError_t x = apply_targets(S, moving, st, what);
            if (x != hipSuccess)
                for (auto &r : moving) fail(*r, MD_ERR_HIP, what + " failed: " + hipGetErrorString(x));
        }
        // one iteration for every request with iterations left, in the same launches
        std::vector<std::shared_ptr<Request>> it_set;
        for (auto &r : S->active)
            if (r->left > 0 && r->status == MD_OK) it_set.push_back(r);
        hipError_t e = hipSuccess;
        bool launched = false;
        if (!it_set.empty()) {
            std::vector<std::pair<int, uint64_t>> key;
            for (auto &r : it_set) key.emplace_back(r->sm->id, r->sm->m->gen);
            int p = cur;
            if (S->keys[cur] != key) {
                // the other plan's last launch was waited for at the previous boundary
                p = 1 - cur;
                std::vector<const MeshDev *> devs;
                std::vector<const Factor *> facs;
                for (auto &r : it_set) {
                    devs.push_back(&r->sm->m->dev);
                    facs.push_back(&r->sm->m->fac);
                }
                md_mesh *m0 = it_set[0]->sm->m;
                S->plans[p].pcg = m0->cfg.solver == MD_SOLVER_PCG;
                S->plans[p].pcg_iters = m0->cfg.pcg_max_iters;
                e = build_plan(S->plans[p], devs, facs, m0->fp64(), m0->cfg.refine_steps, st, false);
                S->keys[p] = e == hipSuccess ? key : std::vector<std::pair<int, uint64_t>>();
            }
            if (e == hipSuccess) {
                enqueue_iteration(S->plans[p], st, nullptr);
                e = hipGetLastError();
            }
            if (e == hipSuccess) {
                launched = true;
                cur = p;
                ++S->seq;
                std::lock_guard<std::mutex> lk(S->mu);
                ++S->stats.iterations;
                S->stats.mesh_iterations += it_set.size();
            } else {
                // this boundary's requests fail; the server goes on with the next ones
                for (auto &r : it_set) fail(*r, MD_ERR_HIP, std::string("launch failed: ") + hipGetErrorString(e));
                S->keys[0].clear();
                S->keys[1].clear();
            }
            if (launched)
                for (auto &r : it_set) r->left--;
        }
        // requests without iterations left: their positions to the pinned buffers
        std::vector<std::shared_ptr<Request>> ending;
        for (auto &r : S->active)
            if (r->left == 0) ending.push_back(r);
        if (!ending.empty()) {
            std::lock_guard<std::mutex> lk(S->mu);
            for (auto &r : ending)
                if (r->status == MD_OK && !(r->pinned = S->take_pinned(r->sm)))
                    fail(*r, MD_ERR_NOMEM, "pinned result buffer");
        }
        // the positions go to the pinned buffers on the copy stream, after this boundary's
        // launch (or, when nothing was launched, after everything queued so far), so the copies
        // overlap the next launches; a mesh is not iterated again before its result is delivered
        if (launched) hipEventRecord(S->ev[S->seq % 2], st);
        bool waited = false;
        for (auto &r : ending) {
            if (r->status == MD_OK) {
                if (!waited) {
                    hipEvent_t after = launched ? S->ev[S->seq % 2] : S->ev_flush;
                    if (!launched) hipEventRecord(after, st);
                    hipStreamWaitEvent(S->copy_stream, after, 0);
                    waited = true;
                }
                md_mesh *m = r->sm->m;
                hipError_t x = hipMemcpyAsync(r->pinned, m->dev.pp, sizeof(d4) * m->n, hipMemcpyDeviceToHost,
                                              S->copy_stream);
                if (x != hipSuccess) fail(*r, MD_ERR_HIP, hipGetErrorString(x));
            }
            r->finish_seq = launched ? S->seq : 0;
            S->finishing.push_back(r);
        }
        if (launched) hipEventRecord(S->evc[S->seq % 2], S->copy_stream);
        S->active.erase(std::remove_if(S->active.begin(), S->active.end(),
                                       [](const std::shared_ptr<Request> &r) { return r->left == 0; }),
                        S->active.end());
        // complete what the previous launch finished (this one keeps the device busy), or all
        // queued work when nothing was launched
        hipError_t err = hipSuccess;
        if (launched) {
            if (S->seq > 1) drain(S, S->seq - 1, err);
            else if (S->seq == 1) {
                // first launch: nothing earlier to wait for, except copies of requests that
                // ended before any launch
                bool early = false;
                for (auto &r : S->finishing) early |= r->finish_seq == 0;
                if (early) drain(S, 0, err);
            }
        } else {
            drain(S, 0, err);
        }
        if (err != hipSuccess) {
            for (auto &r : S->active) fail(*r, MD_ERR_HIP, hipGetErrorString(err));
        }
        ++S->boundaries;
    }
}

}  // namespace

extern "C" int md_server_create(md_ctx *ctx, const md_server_config *cfg, md_server **out) {
    if (!ctx || !out) { set_error("md_server_create: invalid argument"); return MD_ERR_INVALID; }
    std::unique_ptr<md_server> S(new md_server());
    S->ctx = ctx;
    if (cfg) S->cfg = *cfg;
    hipSetDevice(ctx->device);
    if (hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S->copy_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&S->ev[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S->ev[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S->evc[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S->evc[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S->ev_flush, hipEventDisableTiming) != hipSuccess) {
        for (hipEvent_t &x : S->ev)
            if (x) hipEventDestroy(x);
        for (hipEvent_t &x : S->evc)
            if (x) hipEventDestroy(x);
        if (S->ev_flush) hipEventDestroy(S->ev_flush);
        if (S->copy_stream) hipStreamDestroy(S->copy_stream);
        if (S->stream) hipStreamDestroy(S->stream);
        set_error("md_server_create: stream or event creation failed");
        return MD_ERR_HIP;
    }
    ctx_retain(ctx);
    S->worker = std::thread(server_loop, S.get());
    *out = S.release();
    return MD_OK;
}

extern "C" int md_server_add_mesh(md_server *S, md_mesh *m, int *mesh_id) {
    if (!S || !m || !mesh_id) { set_error("md_server_add_mesh: invalid argument"); return MD_ERR_INVALID; }
    if (m->ctx != S->ctx) { set_error("md_server_add_mesh: mesh from another context"); return MD_ERR_INVALID; }
    if (m->servers > 0) { set_error("md_server_add_mesh: the mesh is registered with a server already"); return MD_ERR_STATE; }
    if (m->dirty || !m->analyzed || m->cfg.rest_mode != MD_REST_FIXED) {
        set_error("md_server_add_mesh: the mesh must be prepared (md_mesh_prepare) with MD_REST_FIXED");
        return MD_ERR_STATE;
    }
    hipSetDevice(S->ctx->device);
    if (int rc = ensure_resident(m, "md_server_add_mesh")) return rc;
    const size_t bytes = m->arena.size() + m->larena.size();
    if (S->cfg.hbm_budget_bytes > 0 && bytes > S->cfg.hbm_budget_bytes) {
        set_error("md_server_add_mesh: mesh needs " + std::to_string(bytes) + " device bytes, more than the HBM budget of " +
                  std::to_string(S->cfg.hbm_budget_bytes));
        return MD_ERR_BUDGET;
    }
    // its preparation on the context's stream is complete before the server's stream reads it
    if (hipStreamSynchronize(S->ctx->stream) != hipSuccess) { set_error("md_server_add_mesh: synchronisation failed"); return MD_ERR_HIP; }
    // one pinned result buffer now (pinned allocations take milliseconds: made at a boundary,
    // they would stall the worker); more are made only while earlier results wait for their
    // md_server_wait / md_server_poll
    void *buf = nullptr;
    if (hipHostMalloc(&buf, sizeof(d4) * (size_t)std::max(1, m->n), hipHostMallocDefault) != hipSuccess) {
        set_error("md_server_add_mesh: pinned result buffer: allocation failed");
        return MD_ERR_NOMEM;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->pinned_all.push_back(static_cast<d4 *>(buf));
    if (!S->meshes.empty()) {
        md_mesh *m0 = S->meshes[0]->m;
        if (m0->cfg.precision != m->cfg.precision || m0->cfg.refine_steps != m->cfg.refine_steps ||
            m0->cfg.solver != m->cfg.solver || m0->cfg.pcg_max_iters != m->cfg.pcg_max_iters) {
            set_error("md_server_add_mesh: meshes of one server share precision and refinement steps");
            return MD_ERR_INVALID;
        }
    }
    if (S->cfg.hbm_budget_bytes > 0 && S->resident + bytes > S->cfg.hbm_budget_bytes) {
        // Registration beyond the budget (non-resident meshes, evicted to host memory and
        // restored on demand by the admission planner) is switched off: every served mesh
        // stays resident, so the planner never evicts or restores.
        set_error("md_server_add_mesh: the served meshes would need " + std::to_string(S->resident + bytes) +
                  " device bytes, more than the HBM budget of " + std::to_string(S->cfg.hbm_budget_bytes));
        return MD_ERR_BUDGET;
    }
    S->resident += bytes;
    S->peak_resident = std::max(S->peak_resident, S->resident);
    std::unique_ptr<ServedMesh> sm(new ServedMesh());
    sm->m = m;
    sm->id = (int)S->meshes.size();
    sm->bytes = bytes;
    sm->pinned_free.push_back(static_cast<d4 *>(buf));
    S->meshes.push_back(std::move(sm));
    ++m->servers;
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
        ServedMesh *sm = S->meshes[mesh_id].get();
        md_mesh *m = sm->m;
        for (int i = 0; i < k; ++i)
            if (idx[i] < 0 || idx[i] >= m->n || !m->is_con[idx[i]]) {
                set_error("md_server_submit: vertex " + std::to_string(idx[i]) + " is not a constrained vertex of the mesh");
                return MD_ERR_INVALID;
            }
        r->ticket = S->next_ticket++;
        r->sm = sm;
        // a vertex listed twice takes its last target (setConstraint called twice, arap.h:81-85);
        // the kept targets are distinct, so their scatter has no conflicting writes
        if (sm->stamp.size() != (size_t)m->n) sm->stamp.assign((size_t)m->n, 0);
        if (++sm->stamp_gen == 0) {
            std::fill(sm->stamp.begin(), sm->stamp.end(), 0u);
            sm->stamp_gen = 1;
        }
        r->idx.reserve(k);
        r->xyz.reserve(3 * (size_t)k);
        for (int i = k - 1; i >= 0; --i) {
            if (sm->stamp[idx[i]] == sm->stamp_gen) continue;
            sm->stamp[idx[i]] = sm->stamp_gen;
            r->idx.push_back(idx[i]);
            r->xyz.insert(r->xyz.end(), xyz + 3 * (size_t)i, xyz + 3 * (size_t)i + 3);
        }
        r->iters = r->left = iterations;
        r->t_submit = SClock::now();
        S->queue.push_back(r);
        S->tickets[r->ticket] = r;
        ++S->stats.submitted;
        *ticket = r->ticket;
    }
    S->cv_work.notify_one();
    return MD_OK;
}

// Called with S->mu held: takes the finished request out of the ticket map (null when another
// caller took it first). The permutation of its positions into the original vertex order runs
// afterwards without the lock.
static std::shared_ptr<Request> take_finished(md_server *S, uint64_t ticket) {
    auto it = S->tickets.find(ticket);
    if (it == S->tickets.end()) return nullptr;
    std::shared_ptr<Request> r = it->second;
    S->tickets.erase(it);
    return r;
}

static int deliver(md_server *S, std::shared_ptr<Request> r, double *xyz, double *latency_ms) {
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    int rc = r->status;
    if (rc != MD_OK) set_error("request failed: " + r->error);
    else if (xyz && r->pinned) {
        md_mesh *m = r->sm->m;
        for (int iv = 0; iv < m->n; ++iv) {
            int v = m->topo.original_of[iv];
            xyz[3 * (size_t)v] = r->pinned[iv].x;
            xyz[3 * (size_t)v + 1] = r->pinned[iv].y;
            xyz[3 * (size_t)v + 2] = r->pinned[iv].z;
        }
    }
    if (r->pinned) {
        std::lock_guard<std::mutex> lk(S->mu);
        r->sm->pinned_free.push_back(r->pinned);
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
        if (it == S->tickets.end()) { set_error("md_server_poll: unknown ticket (or already taken)"); return MD_ERR_INVALID; }
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
        if (it == S->tickets.end()) { set_error("md_server_wait: unknown ticket (or already taken)"); return MD_ERR_INVALID; }
        r = it->second;
        bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), [&] { return r->done; });
        if (!ok) { set_error("md_server_wait: timed out"); return MD_ERR_TIMEOUT; }
        // another poll or wait may have taken the ticket while this one slept
        if (!take_finished(S, ticket)) { set_error("md_server_wait: the ticket was taken by another call"); return MD_ERR_INVALID; }
    }
    return deliver(S, r, xyz, latency_ms);
}

extern "C" int md_server_get_stats(md_server *S, md_server_stats *out) {
    if (!S || !out) { set_error("md_server_get_stats: invalid argument"); return MD_ERR_INVALID; }
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->stats;
    out->resident_bytes = S->resident;
    out->peak_resident_bytes = S->peak_resident;
    out->meshes = (int)S->meshes.size();
    out->resident_meshes = 0;
    for (auto &x : S->meshes) out->resident_meshes += x->m->evicted ? 0 : 1;
    out->queued = (int)S->queue.size();
    return MD_OK;
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
    hipStreamSynchronize(S->stream);
    hipStreamSynchronize(S->copy_stream);
    for (d4 *p : S->pinned_all) hipHostFree(p);
    for (TargetEntry *p : S->tgt_host)
        if (p) hipHostFree(p);
    if (S->tgt_dev) hipFree(S->tgt_dev);
    // meshes evicted by the server stay on the host until their next use (md_mesh_* restores)
    for (auto &x : S->meshes) --x->m->servers;
    for (hipEvent_t &x : S->ev) hipEventDestroy(x);
    for (hipEvent_t &x : S->evc) hipEventDestroy(x);
    hipEventDestroy(S->ev_flush);
    hipStreamDestroy(S->copy_stream);
    hipStreamDestroy(S->stream);
    delete S;
    ctx_release(ctx);
    return MD_OK;
}
