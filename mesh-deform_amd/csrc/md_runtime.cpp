// C ABI of the ARAP library: context, mesh (the AsRigidAsPossibleDeformation object of
// inc/deform/arap.h), launch plans and graph replay. Batches and the server are in
// md_serve.cpp.
#include "md_runtime.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <atomic>
#include <mutex>
#include <thread>
#include <type_traits>

namespace md {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }

void ctx_retain(md_ctx *c) { c->refs.fetch_add(1); }

void ctx_release(md_ctx *c) {
    if (c->refs.fetch_sub(1) != 1) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
    delete c;
}

#define MD_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            md::set_error(std::string(#call) + ": " + hipGetErrorString(e_));           \
            return MD_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

#define MD_HIP_E(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) return e_;                                                \
    } while (0)

Plan::~Plan() {
    drop_graph();
    if (d_tables) hipFree(d_tables);
    if (h_tables) hipHostFree(h_tables);
    if (d_bmesh) hipFree(d_bmesh);
}

void Plan::drop_graph() {
    if (graph) hipGraphExecDestroy(graph);
    if (graph_multi) hipGraphExecDestroy(graph_multi);
    graph = graph_multi = nullptr;
}

static int cdiv(int a, int b) { return (a + b - 1) / b; }

hipError_t build_plan(Plan &plan, const std::vector<const MeshDev *> &meshes,
                      const std::vector<const Factor *> &factors, bool fp64, int refine_steps,
                      hipStream_t stream, bool sync, bool block_maps) {
    plan.drop_graph();
    plan.fp64 = fp64;
    plan.refine_steps = std::max(1, refine_steps);
    int na = (int)meshes.size();
    plan.nactive = na;
    plan.nlevels = 0;
    for (const Factor *f : factors) plan.nlevels = std::max(plan.nlevels, f->nlevels);
    int nl = plan.nlevels;
    int ntables = 7 + 2 * nl;
    plan.h_prefix.assign((size_t)ntables * (na + 1), 0);
    plan.h_tbase.assign((size_t)ntables * (na + 1), 0);
    plan.off_fwd.assign(nl, 0);
    plan.off_bwd.assign(nl, 0);
    plan.blk_fwd.assign(nl, 0);
    plan.blk_bwd.assign(nl, 0);
    auto table = [&](int t, int &off, int &total, auto count) {
        off = t * (na + 1);
        int *p = plan.h_prefix.data() + off;
        p[0] = 0;
        for (int k = 0; k < na; ++k) p[k + 1] = p[k] + count(k);
        total = p[na];
    };
    table(0, plan.off_face, plan.blk_face, [&](int k) { return cdiv(meshes[k]->F, kBlock); });
    table(1, plan.off_slice, plan.blk_slice,
          [&](int k) { return cdiv(meshes[k]->nslices, kSlicesPerBlock); });
    table(2, plan.off_resid, plan.blk_resid,
          [&](int k) { return cdiv(meshes[k]->nslices_free, kSlicesPerBlock); });
    table(3, plan.off_ident, plan.blk_ident, [&](int k) { return cdiv(meshes[k]->n, kBlock); });
    table(4, plan.off_apply, plan.blk_apply, [&](int k) { return cdiv(meshes[k]->nf, kBlock); });
    table(5, plan.off_sub, plan.blk_sub, [&](int k) { return factors[k]->nsub; });
    table(6, plan.off_subb, plan.blk_subb,
          [&](int k) { return factors[k]->nsub + cdiv(meshes[k]->nlrows, kTile); });
    plan.sub_lds_fwd = 0;
    plan.sub_lds_bwd = 0;
    plan.sub_nodes = 0;
    for (const Factor *f : factors) {
        plan.sub_lds_fwd = std::max(plan.sub_lds_fwd, f->sub_lds_fwd);
        plan.sub_lds_bwd = std::max(plan.sub_lds_bwd, f->sub_lds_bwd);
        plan.sub_nodes = std::max(plan.sub_nodes, f->sub_nodes);
    }
    for (int l = 0; l < nl; ++l) {
        table(7 + 2 * l, plan.off_fwd[l], plan.blk_fwd[l], [&](int k) {
            const Factor *f = factors[k];
            return l < f->nlevels ? f->ft_ptr[l + 1] - f->ft_ptr[l] : 0;
        });
        table(8 + 2 * l, plan.off_bwd[l], plan.blk_bwd[l], [&](int k) {
            const Factor *f = factors[k];
            return l < f->nlevels ? f->bt_ptr[l + 1] - f->bt_ptr[l] : 0;
        });
        for (int k = 0; k < na; ++k) {
            const Factor *f = factors[k];
            plan.h_tbase[(size_t)plan.off_fwd[l] + k] = l < f->nlevels ? f->ft_ptr[l] : 0;
            plan.h_tbase[(size_t)plan.off_bwd[l] + k] = l < f->nlevels ? f->bt_ptr[l] : 0;
        }
    }
    // block -> mesh maps: one entry per block of every table, so kernels of large batches
    // find their mesh with one load instead of a binary search over the prefix table
    plan.bmesh_off.assign(ntables, (size_t)-1);
    plan.h_bmesh.clear();
    if (block_maps && na > 1) {
        for (int t = 0; t < ntables; ++t) {
            const int *p = plan.h_prefix.data() + (size_t)t * (na + 1);
            plan.bmesh_off[t] = plan.h_bmesh.size();
            for (int k = 0; k < na; ++k)
                for (int b = p[k]; b < p[k + 1]; ++b) plan.h_bmesh.push_back(k);
        }
        if (plan.h_bmesh.size() > plan.cap_bmesh) {
            if (plan.d_bmesh) hipFree(plan.d_bmesh);
            plan.d_bmesh = nullptr;
            MD_HIP_E(hipMalloc(&plan.d_bmesh, sizeof(int) * std::max<size_t>(1, plan.h_bmesh.size())));
            plan.cap_bmesh = plan.h_bmesh.size();
        }
        if (!plan.h_bmesh.empty())
            MD_HIP_E(hipMemcpyAsync(plan.d_bmesh, plan.h_bmesh.data(), sizeof(int) * plan.h_bmesh.size(),
                                    hipMemcpyHostToDevice, stream));
    }
    plan.h_meshes.resize(na);
    for (int k = 0; k < na; ++k) plan.h_meshes[k] = *meshes[k];
    // the mesh records, prefix tables and tile bases in one buffer: staged in pinned host
    // memory and uploaded with one asynchronous copy (the mirror is rewritten only after the
    // stream has passed the previous upload: the callers synchronise or alternate two plans)
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_pre = up(sizeof(MeshDev) * (size_t)na), o_tb = o_pre + up(sizeof(int) * plan.h_prefix.size());
    const size_t total = std::max<size_t>(256, o_tb + up(sizeof(int) * plan.h_tbase.size()));
    if (total > plan.cap_tables) {
        if (plan.d_tables) hipFree(plan.d_tables);
        if (plan.h_tables) hipHostFree(plan.h_tables);
        plan.d_tables = plan.h_tables = nullptr;
        plan.cap_tables = 0;
        const size_t cap = total + total / 2;
        MD_HIP_E(hipMalloc((void **)&plan.d_tables, cap));
        MD_HIP_E(hipHostMalloc((void **)&plan.h_tables, cap, hipHostMallocDefault));
        plan.cap_tables = cap;
    }
    std::memcpy(plan.h_tables, plan.h_meshes.data(), sizeof(MeshDev) * (size_t)na);
    std::memcpy(plan.h_tables + o_pre, plan.h_prefix.data(), sizeof(int) * plan.h_prefix.size());
    std::memcpy(plan.h_tables + o_tb, plan.h_tbase.data(), sizeof(int) * plan.h_tbase.size());
    plan.d_meshes = reinterpret_cast<MeshDev *>(plan.d_tables);
    plan.d_prefix = reinterpret_cast<int *>(plan.d_tables + o_pre);
    plan.d_tbase = reinterpret_cast<int *>(plan.d_tables + o_tb);
    MD_HIP_E(hipMemcpyAsync(plan.d_tables, plan.h_tables, total, hipMemcpyHostToDevice, stream));
    return sync ? hipStreamSynchronize(stream) : hipSuccess;
}

// One-mesh plans with at most this many bottom subtrees form the residual inside the subtree
// forward launch: its registers allow 3 waves per SIMD, and 3 x 1024 SIMDs of a 256-CU device
// then hold every subtree at once (more subtrees would queue behind the longer waves).
constexpr int kFusedResidualMaxSubtrees = 3072;

void enqueue_iteration(const Plan &plan, hipStream_t stream, hipEvent_t *ev) {
    if (ev) hipEventRecord(ev[0], stream);
    launch_local(stream, plan.set(plan.off_slice), plan.blk_slice, plan.fp64);
    if (ev) hipEventRecord(ev[1], stream);
    for (int r = 0; r < plan.refine_steps; ++r) {
        if (plan.pcg) {
            launch_residual(stream, plan.set(plan.off_resid), plan.blk_resid, plan.fp64);
            if (ev && r == 0) hipEventRecord(ev[2], stream);
            // Jacobi-PCG on the correction (Chronopoulos-Gear form): init, one launch per CG
            // iteration, apply
            launch_pcg_init(stream, plan.set(plan.off_resid), plan.blk_resid, plan.fp64);
            for (int k = 0; k < plan.pcg_iters; ++k)
                launch_pcg_step(stream, plan.set(plan.off_resid), plan.blk_resid, k, plan.fp64);
            if (ev && r == 0) hipEventRecord(ev[3], stream);
            launch_pcg_apply(stream, plan.set(plan.off_apply), plan.blk_apply, plan.fp64);
            if (ev && r == 0) hipEventRecord(ev[4], stream);
            continue;
        }
        // one-mesh plans form the residual inside the subtree forward launch (subtree rows in
        // LDS, level rows by extra blocks): its phase is then empty and the forward phase
        // holds it; batches and server mixes keep k_residual (occupancy of the subtree sweep)
        const bool fused = plan.nactive == 1 && plan.blk_sub <= kFusedResidualMaxSubtrees;
        if (!fused) launch_residual(stream, plan.set(plan.off_resid), plan.blk_resid, plan.fp64);
        if (ev && r == 0) hipEventRecord(ev[2], stream);
        if (fused)
            launch_sub_fwd(stream, plan.set(plan.off_subb), plan.blk_subb, plan.sub_nodes, plan.sub_lds_fwd, true,
                           plan.fp64);
        else
            launch_sub_fwd(stream, plan.set(plan.off_sub), plan.blk_sub, plan.sub_nodes, plan.sub_lds_fwd, false,
                           plan.fp64);
        for (int l = 0; l < plan.nlevels; ++l)
            launch_fwd(stream, plan.level_set(plan.off_fwd[l], true), plan.blk_fwd[l], l, plan.fp64);
        if (ev && r == 0) hipEventRecord(ev[3], stream);
        for (int l = plan.nlevels - 1; l >= 0; --l)
            launch_bwd(stream, plan.level_set(plan.off_bwd[l], false), plan.blk_bwd[l], l, plan.fp64);
        // subtrees, then p' += x (subtree rows inside their blocks, level rows in extra blocks)
        launch_sub_bwd(stream, plan.set(plan.off_subb), plan.blk_subb, plan.sub_nodes, plan.sub_lds_bwd, plan.fp64);
        if (ev && r == 0) hipEventRecord(ev[4], stream);
    }
}

// Graph replays of one iteration each leave a few microseconds between launches (about 6 us on
// a 0.16 ms iteration); a graph of kGraphIterations iterations takes the bulk of a longer run.
constexpr int kGraphIterations = 8;

static hipError_t capture_iterations(Plan &plan, hipStream_t stream, int count, hipGraphExec_t *out) {
    hipGraph_t g = nullptr;
    MD_HIP_E(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < count; ++i) enqueue_iteration(plan, stream, nullptr);
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) *out = nullptr;
    return e;
}

hipError_t run_iterations(Plan &plan, hipStream_t stream, int iters, bool use_graph) {
    if (iters <= 0) return hipSuccess;
    if (!use_graph) {
        for (int i = 0; i < iters; ++i) enqueue_iteration(plan, stream, nullptr);
        return hipGetLastError();
    }
    // a PCG iteration is already hundreds of launches: one-iteration graphs only
    const bool multi = !plan.pcg;
    if (!plan.graph || (multi && !plan.graph_multi)) {  // both at the first replay: a warm-up call captures them
        plan.drop_graph();
        MD_HIP_E(capture_iterations(plan, stream, 1, &plan.graph));
        if (multi) MD_HIP_E(capture_iterations(plan, stream, kGraphIterations, &plan.graph_multi));
    }
    if (multi)
        for (; iters >= kGraphIterations; iters -= kGraphIterations) MD_HIP_E(hipGraphLaunch(plan.graph_multi, stream));
    for (; iters > 0; --iters) MD_HIP_E(hipGraphLaunch(plan.graph, stream));
    return hipGetLastError();
}

}  // namespace md

using namespace md;

// ------------------------------------------------------------------------------------------
// library
// ------------------------------------------------------------------------------------------

extern "C" void md_config_default(md_config *cfg) {
    if (!cfg) return;
    cfg->precision = MD_PREC_FP64;
    cfg->solver = MD_SOLVER_LDLT;
    cfg->refine_steps = 1;
    cfg->rest_mode = MD_REST_REFERENCE;
    cfg->pcg_max_iters = 200;
    cfg->pcg_tol = 1e-7;
    cfg->use_graph = 1;
    cfg->subtree_budget = kSubBudgetDefault;
    cfg->amalg_ratio = kAmalgRatioDefault;
    cfg->factor_device = 1;
    cfg->level_pairs = 1;
}

extern "C" const char *md_last_error(void) { return g_err.c_str(); }

extern "C" const char *md_version(void) { return "mdarap 0.1 (gfx950)"; }

extern "C" int md_device_count(int *count) {
    if (!count) { set_error("md_device_count: null pointer"); return MD_ERR_INVALID; }
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return MD_OK;
}

extern "C" int md_ctx_create(int device, md_ctx **out) {
    if (!out) { set_error("md_ctx_create: null pointer"); return MD_ERR_INVALID; }
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || device < 0 || device >= c) {
        set_error("md_ctx_create: no GPU with index " + std::to_string(device));
        return MD_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    MD_HIP(hipGetDeviceProperties(&prop, device));
    std::string arch = prop.gcnArchName;
    if (arch.rfind("gfx950", 0) != 0) {
        set_error("md_ctx_create: device " + std::to_string(device) + " is " + arch +
                  "; this library is built for gfx950 (MI355X) only");
        return MD_ERR_NO_DEVICE;
    }
    MD_HIP(hipSetDevice(device));
    std::unique_ptr<md_ctx> ctx(new md_ctx());
    ctx->device = device;
    ctx->arch = arch;
    MD_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // the static workspace of the non-allocating entry points, at its address on this device
    if (hipError_t e = scratch_address(&ctx->scratch); e != hipSuccess || !ctx->scratch) {
        hipStreamDestroy(ctx->stream);
        set_error(std::string("md_ctx_create: address of the device workspace: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    *out = ctx.release();
    return MD_OK;
}

extern "C" int md_ctx_destroy(md_ctx *ctx) {
    if (!ctx) return MD_OK;
    ctx_release(ctx);
    return MD_OK;
}

extern "C" int md_ctx_synchronize(md_ctx *ctx) {
    if (!ctx) { set_error("md_ctx_synchronize: null context"); return MD_ERR_INVALID; }
    MD_HIP(hipStreamSynchronize(ctx->stream));
    return MD_OK;
}

// ------------------------------------------------------------------------------------------
// mesh
// ------------------------------------------------------------------------------------------

extern "C" int md_mesh_create(md_ctx *ctx, int n, int F, const int *tri, const md_config *cfg,
                              md_mesh **out) {
    if (!ctx || !out || n < 0 || F < 0 || (F > 0 && !tri)) {
        set_error("md_mesh_create: invalid argument");
        return MD_ERR_INVALID;
    }
    *out = nullptr;
    for (int64_t k = 0; k < 3 * (int64_t)F; ++k)
        if (tri[k] < 0 || tri[k] >= n) {
            set_error("md_mesh_create: face " + std::to_string(k / 3) + " references vertex " +
                      std::to_string(tri[k]) + " outside [0, " + std::to_string(n) + ")");
            return MD_ERR_INVALID;
        }
    std::unique_ptr<md_mesh> m(new md_mesh());
    m->ctx = ctx;
    if (cfg) m->cfg = *cfg; else md_config_default(&m->cfg);
    if (m->cfg.precision != MD_PREC_FP64 && m->cfg.precision != MD_PREC_FP32_REFINE) {
        set_error("md_mesh_create: unknown precision");
        return MD_ERR_INVALID;
    }
    if (m->cfg.solver != MD_SOLVER_LDLT && m->cfg.solver != MD_SOLVER_PCG) {
        set_error("md_mesh_create: unknown solver");
        return MD_ERR_INVALID;
    }
    if (m->cfg.solver == MD_SOLVER_PCG && (m->cfg.pcg_max_iters < 1 || !(m->cfg.pcg_tol >= 0.0))) {
        set_error("md_mesh_create: PCG needs pcg_max_iters >= 1 and pcg_tol >= 0");
        return MD_ERR_INVALID;
    }
    if (!(m->cfg.amalg_ratio >= 0.0 && m->cfg.amalg_ratio <= 1.0)) {
        set_error("md_mesh_create: amalg_ratio must lie in [0, 1]");
        return MD_ERR_INVALID;
    }
    if (m->cfg.subtree_budget < 0 || m->cfg.subtree_budget > kSubBudgetMax) {
        set_error("md_mesh_create: subtree_budget must lie in [0, " + std::to_string(kSubBudgetMax) + "]");
        return MD_ERR_INVALID;
    }
    m->cfg.refine_steps = std::max(1, m->cfg.refine_steps);
    m->n = n;
    m->F = F;
    m->tri.assign(tri, tri + 3 * (size_t)F);
    m->is_con.assign(n, 0);
    m->con.assign(3 * (size_t)n, 0.0);
    ctx_retain(ctx);
    *out = m.release();
    return MD_OK;
}

extern "C" int md_mesh_destroy(md_mesh *m) {
    if (!m) return MD_OK;
    if (m->servers > 0) {
        set_error("md_mesh_destroy: the mesh is registered with a server (destroy the server first)");
        return MD_ERR_STATE;
    }
    md_ctx *ctx = m->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    delete m;
    ctx_release(ctx);
    return MD_OK;
}

extern "C" int md_mesh_set_constraint(md_mesh *m, int vidx, const double *xyz) {
    if (!m || !xyz) { set_error("md_mesh_set_constraint: null pointer"); return MD_ERR_INVALID; }
    if (m->servers > 0) {
        set_error("md_mesh_set_constraint: the mesh is registered with a server (drags go through md_server_submit)");
        return MD_ERR_STATE;
    }
    if (vidx < 0 || vidx >= m->n) {
        set_error("md_mesh_set_constraint: vertex index " + std::to_string(vidx) +
                  " outside [0, " + std::to_string(m->n) + ")");
        return MD_ERR_INVALID;
    }
    if (!m->is_con[vidx]) { m->is_con[vidx] = 1; m->ncon++; m->con_set_changed = true; }
    std::memcpy(&m->con[3 * (size_t)vidx], xyz, 3 * sizeof(double));
    m->dirty = true;
    return MD_OK;
}

extern "C" int md_mesh_set_constraints(md_mesh *m, int k, const int *idx, const double *xyz) {
    if (!m || k < 0 || (k > 0 && (!idx || !xyz))) {
        set_error("md_mesh_set_constraints: invalid argument");
        return MD_ERR_INVALID;
    }
    if (m->servers > 0) {
        set_error("md_mesh_set_constraints: the mesh is registered with a server (drags go through md_server_submit)");
        return MD_ERR_STATE;
    }
    for (int i = 0; i < k; ++i)
        if (idx[i] < 0 || idx[i] >= m->n) {
            set_error("md_mesh_set_constraints: vertex index " + std::to_string(idx[i]) +
                      " outside [0, " + std::to_string(m->n) + ")");
            return MD_ERR_INVALID;
        }
    for (int i = 0; i < k; ++i) md_mesh_set_constraint(m, idx[i], xyz + 3 * (size_t)i);
    return MD_OK;
}

namespace {

using Clock = std::chrono::steady_clock;

// Host analysis for the current constrained set: ordering, topology, symbolic factor and
// the device arena with every static array uploaded. Positions come from m->rest.
// Host part of the analysis: ordering, topology, free-free pattern, symbolic factor.
void analyze_host(md_mesh *m) {
    int n = m->n;
    const bool pairs = m->cfg.level_pairs != 0 && m->cfg.solver == MD_SOLVER_LDLT;
    std::vector<int> group;
    std::vector<int> perm = order_free_vertices(n, m->F, m->tri.data(), m->rest.data(), m->is_con,
                                                pairs ? &group : nullptr, m->cfg.amalg_ratio);
    m->topo = Topology();
    build_topology(n, m->F, m->tri.data(), perm, m->is_con, m->topo);
    const Topology &t = m->topo;
    int nf = t.nf;
    // free-free pattern in solve order, from the slots (ascending original neighbour id)
    m->adj_ptr.assign(nf + 1, 0);
    m->adj.clear();
    m->adj_slot.clear();
    for (int i = 0; i < nf; ++i) {
        int s = i / kSlice, l = i % kSlice;
        for (int q = t.slice_ptr[s]; q < t.slice_ptr[s + 1]; ++q) {
            int64_t g = (int64_t)q * kSlice + l;
            int j = t.sell_col[g];
            if (j == i) continue;
            if (j < nf) { m->adj.push_back(j); m->adj_slot.push_back(g); }
        }
        m->adj_ptr[i + 1] = (int)m->adj.size();
    }
    m->fac = Factor();
    if (nf > 0 && nf < n && m->cfg.solver == MD_SOLVER_LDLT)
        symbolic(nf, m->adj_ptr, m->adj, m->fac, m->cfg.subtree_budget, m->cfg.amalg_ratio, pairs ? &group : nullptr);
    else m->fac.n = nf;
}

int upload_layout(md_mesh *m);

// Device part of the analysis: the arena with every static array uploaded, and the plan.
int analyze_device(md_mesh *m) {
    const int n = m->n;
    const Topology &t = m->topo;
    const int nf = t.nf;
    const Factor &f = m->fac;
    // arena
    bool fp64 = m->fp64();
    size_t vsz = fp64 ? sizeof(d4) : sizeof(f4);
    size_t psz = fp64 ? sizeof(double) : sizeof(float);
    int64_t slots = t.total_slots();
    Arena &A = m->arena;
    A.reset();
    size_t o_slice = A.reserve(sizeof(int) * (t.nslices + 1));
    size_t o_col = A.reserve(sizeof(int) * slots);
    size_t o_a32 = A.reserve(sizeof(float) * 3 * slots);
    size_t o_a64 = A.reserve(sizeof(double) * 3 * slots);
    size_t o_w64 = A.reserve(sizeof(double) * slots);
    size_t o_tri = A.reserve(sizeof(int) * 3 * (size_t)m->F);
    size_t o_fe = A.reserve(sizeof(double) * 3 * (size_t)m->F);
    size_t o_sfp = A.reserve(sizeof(int) * (slots + 1));
    size_t o_sf = A.reserve(sizeof(int) * std::max<size_t>(1, t.slot_fe.size()));
    size_t o_p = A.reserve(sizeof(d4) * n);
    size_t o_pp = A.reserve(sizeof(d4) * n);
    size_t o_q = A.reserve(vsz * n);
    size_t o_r = A.reserve(vsz * std::max(1, nf));
    size_t o_z = A.reserve(vsz * std::max(1, nf));
    int ns = f.ns;
    auto last = [](const std::vector<int64_t> &v) { return v.empty() ? (int64_t)0 : v.back(); };
    size_t o_cp = A.reserve(sizeof(int) * (ns + 1));
    size_t o_ch = A.reserve(sizeof(int) * std::max<size_t>(1, f.child.size()));
    size_t o_pc = A.reserve(psz * std::max<int64_t>(1, last(f.pc_off)));
    size_t o_pr = A.reserve(psz * std::max<int64_t>(1, last(f.pr_off)));
    size_t o_fac = A.reserve(sizeof(FacDev) * std::max(1, ns));
    size_t o_hl = A.reserve(sizeof(int) * std::max(1, ns));
    size_t o_or = A.reserve(sizeof(int) * std::max<size_t>(1, f.offrows.size()));
    size_t o_rel = A.reserve(sizeof(int) * std::max<size_t>(1, f.rel.size()));
    int eblocks = std::max(1, cdiv(t.nslices, kSlicesPerBlock));
    size_t o_en = A.reserve(sizeof(double) * eblocks);
    bool pcg = m->cfg.solver == MD_SOLVER_PCG;
    int nblk = std::max(1, cdiv(nf, kBlock));
    size_t ssz = fp64 ? sizeof(double) : sizeof(float);
    size_t o_w32 = A.reserve(sizeof(float) * slots);
    size_t o_dg = A.reserve(ssz * std::max(1, pcg ? nf : 1));
    size_t o_cx = A.reserve(vsz * std::max(1, pcg ? nf : 1));
    size_t o_cpp = A.reserve(vsz * std::max(1, pcg ? nf : 1));
    size_t o_cv = A.reserve(vsz * (pcg ? 5 * (size_t)std::max(1, nf) : 1));
    size_t o_cpt = A.reserve(vsz * (pcg ? 4 * (size_t)nblk + 5 : 1));
    size_t o_tx = A.reserve(sizeof(double) * 3 * std::max(1, n - nf));
    size_t o_orig = A.reserve(sizeof(int) * std::max(1, n));
    size_t o_xyz = A.reserve(sizeof(double) * 3 * std::max(1, n));
    if (A.allocate() != hipSuccess) {
        set_error("device allocation of " + std::to_string(A.size()) + " bytes failed");
        A.reset();
        return MD_ERR_NOMEM;
    }
    MeshDev &d = m->dev;
    d = MeshDev{};
    d.n = n; d.nf = nf; d.nslices = t.nslices; d.ns = ns; d.F = m->F;
    d.nslices_free = cdiv(nf, kSlice);
    d.slots = slots;
    d.slice_ptr = (const int *)A.at(o_slice);
    d.sell_col = (const int *)A.at(o_col);
    d.a32 = (float *)A.at(o_a32);
    d.a64 = (double *)A.at(o_a64);
    d.w64 = (double *)A.at(o_w64);
    d.tri = (const int *)A.at(o_tri);
    d.fe_val = (double *)A.at(o_fe);
    d.slot_fe_ptr = (const int *)A.at(o_sfp);
    d.slot_fe = (const int *)A.at(o_sf);
    d.p = (d4 *)A.at(o_p);
    d.pp = (d4 *)A.at(o_pp);
    d.quat = A.at(o_q);
    d.r = A.at(o_r); d.z = A.at(o_z);
    d.child_ptr = (const int *)A.at(o_cp); d.child = (const int *)A.at(o_ch);
    d.pc = A.at(o_pc); d.pr = A.at(o_pr);
    m->d_fac = (const FacDev *)A.at(o_fac);
    m->d_hlist = (const int *)A.at(o_hl);
    m->d_offrows = (const int *)A.at(o_or);
    m->d_rel = (const int *)A.at(o_rel);
    d.energy_part = (double *)A.at(o_en);
    d.w32 = (float *)A.at(o_w32);
    d.diag = A.at(o_dg);
    d.cg_x = A.at(o_cx); d.cg_p = A.at(o_cpp); d.cg_v = A.at(o_cv); d.cg_part = A.at(o_cpt);
    d.nblk_free = nblk;
    d.pcg_tol2 = m->cfg.pcg_tol * m->cfg.pcg_tol;
    m->d_tgt_xyz = (double *)A.at(o_tx);
    m->d_orig = (const int *)A.at(o_orig);
    m->d_xyz = (double *)A.at(o_xyz);
    hipStream_t st = m->ctx->stream;
    auto up = [&](const void *dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice, st);
    };
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chk(up(d.slice_ptr, t.slice_ptr.data(), sizeof(int) * (t.nslices + 1)));
    chk(up(d.sell_col, t.sell_col.data(), sizeof(int) * slots));
    chk(up(d.tri, t.tri_internal.data(), sizeof(int) * 3 * (size_t)m->F));
    chk(up(d.slot_fe_ptr, t.slot_fe_ptr.data(), sizeof(int) * (slots + 1)));
    chk(up(d.slot_fe, t.slot_fe.data(), sizeof(int) * t.slot_fe.size()));
    chk(up(m->d_orig, t.original_of.data(), sizeof(int) * n));
    std::vector<FacDev> fh(ns);
    for (int k = 0; k < ns; ++k) {
        FacDev &x = fh[k];
        x.c0 = f.c0[k]; x.w = f.w[k]; x.noff = f.noff[k]; x.roff = f.roff[k];
        x.cb = f.child_ptr[k]; x.ce = f.child_ptr[k + 1]; x.m = f.w[k] + f.noff[k]; x.pad = 0;
        x.poff = f.poff[k]; x.soff = f.woff_s[k]; x.toff = f.woff_t[k]; x.uoff = f.woff_u[k];
        x.pc_off = f.pc_off[k]; x.pr_off = f.pr_off[k];
    }
    if (ns > 0) {
        chk(up(d.child_ptr, f.child_ptr.data(), sizeof(int) * (ns + 1)));
        chk(up(d.child, f.child.data(), sizeof(int) * f.child.size()));
        chk(up(m->d_fac, fh.data(), sizeof(FacDev) * ns));
        chk(up(m->d_hlist, f.hlist.data(), sizeof(int) * ns));
        chk(up(m->d_offrows, f.offrows.data(), sizeof(int) * f.offrows.size()));
        chk(up(m->d_rel, f.rel.data(), sizeof(int) * f.rel.size()));
    }
    chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        set_error(std::string("upload of the mesh analysis failed: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    int rc = upload_layout(m);
    if (rc) return rc;
    ++m->gen;
    m->analyzed = true;
    return MD_OK;
}

// The sweep layout of the factor (device_layout) in its own arena, and the mesh's plan.
// Called after the analysis and again when the layout is re-chosen (relayout); the panels
// and every other device array are untouched.
int upload_layout(md_mesh *m) {
    const Factor &f = m->fac;
    const int ns = f.ns;
    const bool fp64 = m->fp64();
    const size_t vsz = fp64 ? sizeof(d4) : sizeof(f4);
    auto last = [](const std::vector<int64_t> &v) { return v.empty() ? (int64_t)0 : v.back(); };
    Arena &A = m->larena;
    A.reset();
    size_t o_up = A.reserve(vsz * std::max<int64_t>(1, last(f.ubase)));
    size_t o_xp = A.reserve(vsz * std::max<int64_t>(1, last(f.xoff)));
    const size_t nft = f.ft.size() / 4, nbt = f.bt.size() / 4;
    size_t o_udst = A.reserve(sizeof(int64_t) * std::max<size_t>(1, f.udst.size()));
    size_t o_ri = A.reserve(sizeof(int) * std::max<size_t>(4, f.rowinfo.size()));
    size_t o_xl = A.reserve(sizeof(int) * std::max<size_t>(4, f.xloc.size()));
    size_t o_ftile = A.reserve(sizeof(TileDev) * std::max<size_t>(1, nft));
    size_t o_btile = A.reserve(sizeof(TileDev) * std::max<size_t>(1, nbt));
    size_t o_sub = A.reserve(sizeof(int) * std::max<size_t>(4, f.sub.size()));
    size_t o_subh = A.reserve(sizeof(int) * std::max<size_t>(1, f.sub_h.size()));
    size_t o_snode = A.reserve(sizeof(int) * std::max<size_t>(16, f.snode.size()));
    size_t o_scp = A.reserve(sizeof(int) * std::max<size_t>(2, f.scp.size()));
    size_t o_sfc = A.reserve(sizeof(int) * std::max<size_t>(1, f.sfc.size()));
    size_t o_sbo = A.reserve(sizeof(int) * std::max<size_t>(2, f.sbo.size()));
    std::vector<int> lrows;  // rows of level-swept supernodes
    for (int k = 0; k < ns; ++k)
        if (f.level[k] >= 0)
            for (int q = 0; q < f.w[k]; ++q) lrows.push_back(f.c0[k] + q);
    size_t o_lr = A.reserve(sizeof(int) * std::max<size_t>(1, lrows.size()));
    size_t o_srow = A.reserve(sizeof(int) * std::max<size_t>(2, f.srow.size()));
    size_t o_sfi = A.reserve(sizeof(int) * std::max<size_t>(1, f.sfi.size()));
    if (A.allocate() != hipSuccess) {
        set_error("device allocation of " + std::to_string(A.size()) + " bytes (sweep layout) failed");
        A.reset();
        return MD_ERR_NOMEM;
    }
    MeshDev &d = m->dev;
    d.upart = A.at(o_up); d.xpart = A.at(o_xp);
    d.udst = (const long long *)A.at(o_udst);
    d.rowinfo = (const int4 *)A.at(o_ri); d.xloc = (const int4 *)A.at(o_xl);
    d.ftile = (const TileDev *)A.at(o_ftile); d.btile = (const TileDev *)A.at(o_btile);
    d.sub = (const int4 *)A.at(o_sub); d.sub_h = (const int *)A.at(o_subh);
    d.snode = (const int4 *)A.at(o_snode); d.scp = (const int2 *)A.at(o_scp);
    d.sfc = (const int *)A.at(o_sfc); d.sbo = (const int2 *)A.at(o_sbo);
    d.nsub = f.nsub;
    d.nlrows = (int)lrows.size();
    d.lrows = (const int *)A.at(o_lr);
    d.srow = (const int2 *)A.at(o_srow);
    d.sfi = (const int *)A.at(o_sfi);
    hipStream_t st = m->ctx->stream;
    auto up = [&](const void *dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice, st);
    };
    auto sn_of = [&](int k) {
        SnDev x;
        x.c0 = f.c0[k]; x.w = f.w[k]; x.noff = f.noff[k]; x.roff = f.roff[k];
        x.kc = f.kc[k]; x.kr = f.kr[k]; x.uk = f.uk[k];
        x.dstride = f.parent[k] >= 0 ? f.w[f.parent[k]] + f.noff[f.parent[k]] : 0;
        x.pc_off = f.pc_off[k]; x.pr_off = f.pr_off[k]; x.ubase = f.ubase[k]; x.xoff = f.xoff[k];
        return x;
    };
    std::vector<TileDev> fth(nft), bth(nbt);
    for (size_t i = 0; i < nft; ++i)
        fth[i] = TileDev{sn_of(f.ft[4 * i]), f.ft[4 * i + 1], f.ft[4 * i + 2], 0, 0};
    for (size_t i = 0; i < nbt; ++i)
        bth[i] = TileDev{sn_of(f.bt[4 * i]), f.bt[4 * i + 1], f.bt[4 * i + 2], 0, 0};
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    if (ns > 0) {
        chk(up(d.udst, f.udst.data(), sizeof(int64_t) * f.udst.size()));
        // slots no child writes must read as zero, every iteration
        chk(hipMemsetAsync(d.upart, 0, vsz * std::max<int64_t>(1, last(f.ubase)), st));
        chk(up(d.rowinfo, f.rowinfo.data(), sizeof(int) * f.rowinfo.size()));
        chk(up(d.xloc, f.xloc.data(), sizeof(int) * f.xloc.size()));
        chk(up(d.ftile, fth.data(), sizeof(TileDev) * nft));
        chk(up(d.btile, bth.data(), sizeof(TileDev) * nbt));
        chk(up(d.sub, f.sub.data(), sizeof(int) * f.sub.size()));
        chk(up(d.sub_h, f.sub_h.data(), sizeof(int) * f.sub_h.size()));
        chk(up(d.snode, f.snode.data(), sizeof(int) * f.snode.size()));
        chk(up(d.scp, f.scp.data(), sizeof(int) * f.scp.size()));
        chk(up(d.sfc, f.sfc.data(), sizeof(int) * f.sfc.size()));
        chk(up(d.sbo, f.sbo.data(), sizeof(int) * f.sbo.size()));
        chk(up(d.lrows, lrows.data(), sizeof(int) * lrows.size()));
        chk(up(d.srow, f.srow.data(), sizeof(int) * f.srow.size()));
        chk(up(d.sfi, f.sfi.data(), sizeof(int) * f.sfi.size()));
    }
    chk(hipStreamSynchronize(st));
    if (e != hipSuccess) {
        set_error(std::string("upload of the sweep layout failed: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    std::vector<const MeshDev *> ms{&m->dev};
    std::vector<const Factor *> fs{&m->fac};
    m->plan.pcg = m->cfg.solver == MD_SOLVER_PCG;
    m->plan.pcg_iters = m->cfg.pcg_max_iters;
    e = build_plan(m->plan, ms, fs, fp64, m->cfg.refine_steps, st);
    if (e != hipSuccess) {
        set_error(std::string("launch plan upload failed: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    return MD_OK;
}

// A copy ordered after the work queued on the mesh's stream (a non-blocking stream, which a
// plain hipMemcpy on the null stream would not wait for), then waited for.
hipError_t copy_on_stream(md_mesh *m, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, m->ctx->stream);
    return e == hipSuccess ? hipStreamSynchronize(m->ctx->stream) : e;
}

// Uploads the caller's positions (original order, 3 per vertex) and spreads them to internal
// order on the device: into p (rest, may be null) and pp (current positions with the targets
// of the constrained vertices pasted, may be null). Stream-ordered, no synchronisation.
// Pinned staging of the packed positions on their way back: downloads of at least kPinMinBytes,
// from a mesh's second one on, in at most kPinChunks chunks of at least kPinChunkMin bytes, each
// chunk copied to the caller's buffer as soon as its DMA has arrived. For 3.9 MB (the 164k
// icosphere) a pageable copy back measured 231 us, a pinned DMA 77 us and the host copy out of
// the pinned buffer 131 us; the chunked form took a frame's download from 0.24 to 0.20 ms. (The
// same staging toward the device made the upload slower: 0.32 against 0.21 ms.)
constexpr size_t kPinMinBytes = size_t(2) << 20, kPinChunkMin = size_t(1) << 20;
static bool pin_ready(md_mesh *m, size_t bytes, int calls) {
    if (bytes < kPinMinBytes || calls < 2) return false;
    if (m->xyz_pin_bytes >= bytes) return true;
    m->release_pin();
    void *h = nullptr;
    if (hipHostMalloc(&h, bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return false;  // the direct copy remains
    }
    m->xyz_pin = static_cast<double *>(h);
    m->xyz_pin_bytes = bytes;
    for (auto &ev : m->pin_ev)
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            m->release_pin();
            return false;
        }
    return true;
}
static size_t pin_chunk(size_t bytes) {
    const size_t c = std::max(kPinChunkMin, (bytes + md_mesh::kPinChunks - 1) / md_mesh::kPinChunks);
    return (c + 7) & ~size_t(7);
}

hipError_t upload_positions(md_mesh *m, const double *xyz, d4 *p, d4 *pp) {
    hipStream_t st = m->ctx->stream;
    const int n = m->n, nf = m->topo.nf;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemcpyAsync(m->d_xyz, xyz, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && pp && n > nf) {
        m->tgt_stage.resize(3 * (size_t)(n - nf));
        for (int iv = nf; iv < n; ++iv)
            std::memcpy(&m->tgt_stage[3 * (size_t)(iv - nf)], &m->con[3 * (size_t)m->topo.original_of[iv]],
                        3 * sizeof(double));
        e = hipMemcpyAsync(m->d_tgt_xyz, m->tgt_stage.data(), sizeof(double) * m->tgt_stage.size(),
                           hipMemcpyHostToDevice, st);
    }
    if (e != hipSuccess) return e;
    launch_unpack_xyz(st, m->d_xyz, m->d_orig, n, nf, m->d_tgt_xyz, p, pp);
    return hipGetLastError();
}

// Current positions to the caller's packed xyz (original order), packed on the device.
hipError_t download_positions(md_mesh *m, double *xyz) {
    hipStream_t st = m->ctx->stream;
    if (m->n == 0) return hipSuccess;
    launch_pack_xyz(st, m->dev.pp, m->d_orig, m->n, m->d_xyz);
    hipError_t e = hipGetLastError();
    const size_t bytes = sizeof(double) * 3 * (size_t)m->n;
    if (e == hipSuccess && pin_ready(m, bytes, ++m->downloads)) {
        // every chunk's DMA queued (each followed by its event), then each chunk copied out as
        // soon as it has arrived
        const size_t ch = pin_chunk(bytes);
        char *pin = reinterpret_cast<char *>(m->xyz_pin), *out = reinterpret_cast<char *>(xyz);
        const char *src = reinterpret_cast<const char *>(m->d_xyz);
        int c = 0;
        for (size_t off = 0; off < bytes && e == hipSuccess; off += ch, ++c) {
            e = hipMemcpyAsync(pin + off, src + off, std::min(ch, bytes - off), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipEventRecord(m->pin_ev[c], st);
        }
        c = 0;
        for (size_t off = 0; off < bytes && e == hipSuccess; off += ch, ++c) {
            e = hipEventSynchronize(m->pin_ev[c]);
            if (e == hipSuccess) std::memcpy(out + off, pin + off, std::min(ch, bytes - off));
        }
        return e;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(xyz, m->d_xyz, bytes, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

hipError_t compute_cotan(md_mesh *m) {
    const Plan &P = m->plan;
    hipStream_t st = m->ctx->stream;
    launch_cotan_face(st, P.set(P.off_face), P.blk_face);
    launch_cotan_slot(st, P.set(P.off_slice), P.blk_slice);
    return hipGetLastError();
}

// FP64 numeric factorization on the host from the device cotan weights; uploads panels.
// With the PCG solver only the Jacobi diagonal is formed (on the device).
int factor_pcg(md_mesh *m) {
    const Plan &P = m->plan;
    launch_pcg_diag(m->ctx->stream, P.set(P.off_resid), P.blk_resid, m->fp64());
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->ctx->stream);
    if (e != hipSuccess) { set_error(std::string("PCG diagonal: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    m->factored = true;
    return MD_OK;
}

// Cotan weights of the device back to the host (input of the host factorization).
int weights_download(md_mesh *m) {
    int64_t slots = m->topo.total_slots();
    m->w_host.resize(slots);
    if (copy_on_stream(m, m->w_host.data(), m->dev.w64, sizeof(double) * slots, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("download of cotan weights failed");
        return MD_ERR_HIP;
    }
    return MD_OK;
}

// FP64 numeric factorization on the host and the two device panel layouts (column-major
// G_s for the forward sweep, row-major [Dinv_s; G_s] for the backward sweep), rounded to the
// solve precision. Thread-safe across meshes. Returns false with err set on failure.
bool factor_host(md_mesh *m, std::string &err) {
    auto t0 = Clock::now();
    const Topology &t = m->topo;
    int nf = t.nf;
    const std::vector<double> &w = m->w_host;
    std::vector<double> diag(nf, 0.0), val(m->adj.size());
    for (int i = 0; i < nf; ++i) {
        int s = i / kSlice, l = i % kSlice;
        double acc = 0.0;  // L_ii sums every neighbour weight in ascending id (arap.h:332)
        for (int q = t.slice_ptr[s]; q < t.slice_ptr[s + 1]; ++q) {
            int64_t g = (int64_t)q * kSlice + l;
            if (t.sell_col[g] != i) acc += w[g];
        }
        diag[i] = acc;
    }
    for (size_t k = 0; k < m->adj.size(); ++k) val[k] = -w[m->adj_slot[k]];
    if (!numeric(m->adj_ptr, m->adj, val, diag, m->fac, err)) {
        m->factored = false;
        return false;
    }
    const Factor &f = m->fac;
    size_t npc = (size_t)f.pc_off[f.ns], npr = (size_t)f.pr_off[f.ns];
    size_t es = m->fp64() ? sizeof(double) : sizeof(float);
    m->pc_host.resize(npc * es);
    m->pr_host.resize(npr * es);
    for (int s = 0; s < f.ns; ++s) {
        int ww = f.w[s], no = f.noff[s], mm = ww + no;
        const double *P = f.panel.data() + f.poff[s];
        for (int q = 0; q < ww; ++q) {
            for (int r = 0; r < no; ++r) {
                size_t k = f.pc_off[s] + (size_t)q * no + r;
                double v = P[(size_t)q * mm + ww + r];
                if (m->fp64()) reinterpret_cast<double *>(m->pc_host.data())[k] = v;
                else reinterpret_cast<float *>(m->pc_host.data())[k] = (float)v;
            }
            for (int p = 0; p < mm; ++p) {
                size_t k = f.pr_off[s] + (size_t)p * ww + q;
                double v = P[(size_t)q * mm + p];
                if (m->fp64()) reinterpret_cast<double *>(m->pr_host.data())[k] = v;
                else reinterpret_cast<float *>(m->pr_host.data())[k] = (float)v;
            }
        }
    }
    m->fac.panel.clear();
    m->fac.panel.shrink_to_fit();
    m->factor_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    m->setup_ms_host += m->factor_ms;
    return true;
}

// FP64 numeric factorization on the device from the device cotan weights, straight into
// the sweeps' panels: diagonal, then per height of the supernodal tree the front assembly and
// the dense partial factorization, then the conversion to the solve precision. The FP64
// workspace lives for the call only.
int factor_device(md_mesh *m) {
    auto t0 = Clock::now();
    const Factor &f = m->fac;
    const int ns = f.ns, nf = m->topo.nf;
    hipStream_t st = m->ctx->stream;
    // per height: workgroups per supernode of the multi-workgroup stages, for the height's largest
    // front (assembly one per asm_cols front columns: 64, halved down to 8 while the height has
    // fewer than 2048 assembly workgroups, since the top fronts take the largest children; the
    // products one per 64 x 64 output tile), and the products' splits over K
    struct HeightLaunch {
        int count, mw, mm, asm_cols;
        FacSplit dinv, g, u;
    };
    const int nh = (int)f.hptr.size() - 1;
    std::vector<HeightLaunch> hl(nh);
    size_t nW = 0;
    for (int h = 0; h < nh; ++h) {
        HeightLaunch &L = hl[h];
        L = HeightLaunch{f.hptr[h + 1] - f.hptr[h], 0, 1, 64, {1, 1}, {0, 1}, {0, 1}};
        auto lower = [](int b) { return b * (b + 1) / 2; };
        for (int i = f.hptr[h]; i < f.hptr[h + 1]; ++i) {
            const int w = f.w[f.hlist[i]], no = f.noff[f.hlist[i]];
            const int bw = (w + 63) / 64, bo = (no + 63) / 64;
            L.mw = std::max(L.mw, w);
            L.mm = std::max(L.mm, w + no);
            L.dinv.per = std::max(L.dinv.per, lower(bw));
            L.g.per = std::max(L.g.per, bo * bw);
            L.u.per = std::max(L.u.per, lower(bo));
        }
        if (L.mm <= kFacSmallFront) continue;  // small fronts only: one fused launch
        while (L.asm_cols > 8 && (long long)L.count * ((L.mm + L.asm_cols - 1) / L.asm_cols) < 2048) L.asm_cols /= 2;
        const int chunks = (L.mw + 31) / 32;
        for (FacSplit *p : {&L.dinv, &L.g, &L.u}) {
            if (p->per == 0) continue;
            p->ks = fac_split_k((long long)L.count * p->per, chunks);
            if (p->ks > 1) nW = std::max(nW, (size_t)fac_split_k_doubles((long long)L.count * p->per * p->ks));
        }
    }
    const size_t nP = (size_t)f.poff[ns], nS = (size_t)f.woff_s[ns], nT = (size_t)f.woff_t[ns],
                 nU = (size_t)f.woff_u[ns];
    const size_t bytes = sizeof(double) * (nP + nS + nT + nU + nW + (size_t)nf) + 2 * sizeof(unsigned long long);
    const bool keep = m->cfg.rest_mode == MD_REST_REFERENCE;  // refactors on every deform call
    if (m->fac_ws && m->fac_ws_bytes != bytes) m->release_fac_ws();
    if (!m->fac_ws) {
        if (hipMalloc(&m->fac_ws, bytes) != hipSuccess) {
            m->fac_ws = nullptr;
            set_error("device factorization workspace of " + std::to_string(bytes) + " bytes: allocation failed");
            return MD_ERR_NOMEM;
        }
        m->fac_ws_bytes = bytes;
    }
    void *ws = m->fac_ws;
    FacArgs A{};
    A.sn = m->d_fac; A.hlist = m->d_hlist; A.offrows = m->d_offrows; A.rel = m->d_rel; A.child = m->dev.child;
    A.slice_ptr = m->dev.slice_ptr; A.sell_col = m->dev.sell_col; A.w64 = m->dev.w64; A.nf = nf;
    double *base = static_cast<double *>(ws);
    A.P = base; A.S = A.P + nP; A.T = A.S + nS; A.U = A.T + nT; A.Wk = A.U + nU; A.diag = A.Wk + nW;
    A.amax = reinterpret_cast<unsigned long long *>(A.diag + nf);
    A.flag = reinterpret_cast<int *>(A.amax + 1);
    A.pc = const_cast<void *>(m->dev.pc);
    A.pr = const_cast<void *>(m->dev.pr);
    A.solve64 = m->fp64() ? 1 : 0;
    hipError_t e = hipMemsetAsync(A.amax, 0, 2 * sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        launch_fac_diag(st, A);
        for (int h = 0; h < nh; ++h) {
            const HeightLaunch &L = hl[h];
            if (L.mm <= kFacSmallFront) {  // small fronts only: one fused launch
                launch_fac_small(st, A, f.hptr[h], L.count, L.mm, L.mw);
                continue;
            }
            launch_fac_height(st, A, f.hptr[h], L.count, (L.mw + 63) / 64, L.asm_cols,
                              (L.mm + L.asm_cols - 1) / L.asm_cols, L.dinv, L.g, L.u);
        }
        e = hipGetLastError();
    }
    int flag = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&flag, A.flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (!keep || e != hipSuccess) m->release_fac_ws();
    if (e != hipSuccess) {
        set_error(std::string("device factorization: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    m->factor_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    if (flag == 1) {
        m->factored = false;
        set_error("factorization failed: a pivot block is not positive definite (a mesh component "
                  "without constraints?)");
        return MD_ERR_FACTOR;
    }
    if (flag != 0) {
        m->factored = false;
        set_error("device factorization: inconsistent symbolic structure");
        return MD_ERR_INVALID;
    }
    m->factored = true;
    return MD_OK;
}

int factor_upload(md_mesh *m) {
    hipError_t e = copy_on_stream(m, const_cast<void *>(m->dev.pc), m->pc_host.data(), m->pc_host.size(),
                                  hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = copy_on_stream(m, const_cast<void *>(m->dev.pr), m->pr_host.data(), m->pr_host.size(),
                           hipMemcpyHostToDevice);
    std::vector<char>().swap(m->pc_host);
    std::vector<char>().swap(m->pr_host);
    if (e != hipSuccess) {
        set_error("panel upload failed");
        return MD_ERR_HIP;
    }
    m->factored = true;
    return MD_OK;
}

// The dirty branch of deform (arap.h:102-120), in phases so that the host work of many
// meshes can run in parallel (md_meshes_prepare) while device work stays on one stream.
struct Prep {
    const double *xyz = nullptr;
    bool full = false, need_analysis = false, need_factor = false, early = false;
};

// Decisions and the rest pose (arap.h:102-108,162-168).
void prep_begin(md_mesh *m, const double *xyz, Prep &p) {
    p.xyz = xyz;
    p.full = m->cfg.rest_mode == MD_REST_REFERENCE || !m->initialized;
    if (p.full) m->rest.assign(xyz, xyz + 3 * (size_t)m->n);
    p.need_analysis = !m->analyzed || m->con_set_changed;
    if (p.need_analysis) m->factored = false;
}

// Device work before the factorization: arena, weights, initial guess, identity rotations.
int prep_device_a(md_mesh *m, Prep &p) {
    if (p.need_analysis) {
        int rc = analyze_device(m);
        if (rc) return rc;
    }
    hipError_t e = hipSuccess;
    // rest pose (when it changes) and the initial guess: the current mesh with the targets
    // pasted (arap.h:162-168, 277-281); in reference rest mode both come from the same upload
    if (p.full) e = upload_positions(m, p.xyz, m->dev.p, m->dev.pp);
    else {
        if (p.need_analysis) e = upload_positions(m, m->rest.data(), m->dev.p, nullptr);
        if (e == hipSuccess) e = upload_positions(m, p.xyz, nullptr, m->dev.pp);
    }
    if (e != hipSuccess) { set_error(std::string("position upload failed: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    if (p.full || p.need_analysis) {
        e = compute_cotan(m);
        if (e != hipSuccess) { set_error(std::string("cotan weights: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    }
    // identity rotations (arap.h:246-249)
    launch_identity_quat(m->ctx->stream, m->plan.set(m->plan.off_ident), m->plan.blk_ident, m->fp64());
    m->initialized = true;
    m->iterated = false;
    // the caller's buffers have been read once this returns (the copies above may be
    // asynchronous), and the cotan weights are complete for the host factorization
    e = hipStreamSynchronize(m->ctx->stream);
    if (e != hipSuccess) { set_error(std::string("position upload failed: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    if (m->ncon == 0) { p.early = true; return MD_OK; }  // arap.h:113-114
    p.need_factor = m->topo.nf > 0 && (p.full || p.need_analysis || !m->factored);
    if (p.need_factor && m->cfg.solver == MD_SOLVER_LDLT && !m->cfg.factor_device) return weights_download(m);
    return MD_OK;
}

// Device work after the factorization.
int prep_device_b(md_mesh *m, Prep &p) {
    if (p.need_factor) {
        int rc = m->cfg.solver == MD_SOLVER_PCG ? factor_pcg(m) : m->cfg.factor_device ? factor_device(m)
                                                                                       : factor_upload(m);
        if (rc) return rc;
    }
    m->dirty = false;
    m->con_set_changed = false;
    return MD_OK;
}

// Returns MD_OK with *early = true when no vertex is constrained (arap.h:113-114).
int setup(md_mesh *m, const double *xyz, bool *early) {
    auto ms_since = [](Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); };
    Prep p;
    auto t0 = Clock::now();
    prep_begin(m, xyz, p);
    m->setup_phase_ms[0] = m->setup_phase_ms[1] = m->setup_phase_ms[2] = 0.0;
    if (p.need_analysis) {
        analyze_host(m);
        m->setup_ms_host = ms_since(t0);
    }
    m->setup_phase_ms[0] = ms_since(t0);
    t0 = Clock::now();
    int rc = prep_device_a(m, p);
    m->setup_phase_ms[1] = ms_since(t0);
    *early = p.early;
    if (rc || p.early) return rc;
    t0 = Clock::now();
    if (p.need_factor && m->cfg.solver == MD_SOLVER_LDLT && !m->cfg.factor_device) {
        std::string err;
        if (!factor_host(m, err)) { set_error(err); return MD_ERR_FACTOR; }
    }
    rc = prep_device_b(m, p);
    m->setup_phase_ms[2] = ms_since(t0);
    return rc;
}

// Runs fn(k) for k in [0, count) on up to nthreads threads.
template <class Fn> void parallel_for(int count, int nthreads, Fn fn) {
    int nt = std::max(1, std::min(nthreads, count));
    if (nt == 1) {
        for (int k = 0; k < count; ++k) fn(k);
        return;
    }
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int k = next++; k < count; k = next++) fn(k);
        });
    for (auto &x : th) x.join();
}

int iterate(md_mesh *m, int iterations) {
    if (iterations <= 0) return MD_OK;
    hipError_t e = run_iterations(m->plan, m->ctx->stream, iterations, m->cfg.use_graph != 0);
    if (e == hipSuccess) e = hipStreamSynchronize(m->ctx->stream);
    if (e != hipSuccess) { set_error(std::string("iteration failed: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    m->iterated = true;
    return MD_OK;
}

}  // namespace

namespace md {

int relayout(md_mesh *m, int budget) {
    hipError_t e = hipStreamSynchronize(m->ctx->stream);  // the old layout may still be in use
    if (e != hipSuccess) { set_error(std::string("relayout: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    m->cfg.subtree_budget = budget;
    device_layout(m->fac, budget);
    int rc = upload_layout(m);
    ++m->gen;
    return rc;
}

StateSnapshot::~StateSnapshot() {
    for (void *p : buf)
        if (p) hipFree(p);
}

hipError_t snapshot_state(const std::vector<md_mesh *> &meshes, hipStream_t st, StateSnapshot &s) {
    for (md_mesh *m : meshes) {
        size_t qb = (m->fp64() ? sizeof(d4) : sizeof(f4)) * (size_t)m->n, pb = sizeof(d4) * (size_t)m->n;
        void *p = nullptr;
        MD_HIP_E(hipMalloc(&p, pb + qb));
        s.buf.push_back(p);
        MD_HIP_E(hipMemcpyAsync(p, m->dev.pp, pb, hipMemcpyDeviceToDevice, st));
        MD_HIP_E(hipMemcpyAsync(static_cast<char *>(p) + pb, m->dev.quat, qb, hipMemcpyDeviceToDevice, st));
    }
    return hipStreamSynchronize(st);
}

hipError_t restore_state(const std::vector<md_mesh *> &meshes, hipStream_t st, const StateSnapshot &s) {
    for (size_t k = 0; k < meshes.size(); ++k) {
        md_mesh *m = meshes[k];
        size_t qb = (m->fp64() ? sizeof(d4) : sizeof(f4)) * (size_t)m->n, pb = sizeof(d4) * (size_t)m->n;
        MD_HIP_E(hipMemcpyAsync(m->dev.pp, s.buf[k], pb, hipMemcpyDeviceToDevice, st));
        MD_HIP_E(hipMemcpyAsync(m->dev.quat, static_cast<const char *>(s.buf[k]) + pb, qb, hipMemcpyDeviceToDevice, st));
    }
    return hipStreamSynchronize(st);
}

hipError_t time_plan(Plan &plan, hipStream_t st, int iters, bool use_graph, double &ms) {
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, st);
    if (e == hipSuccess) e = run_iterations(plan, st, iters, use_graph);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) hipEventDestroy(a);
    if (b) hipEventDestroy(b);
    ms = t;
    return e;
}

// ---- residency: eviction to a host image and restoration -----------------------------------

void HostImage::release() {
    if (a) hipHostFree(a);
    if (l) hipHostFree(l);
    a = l = nullptr;
    asz = lsz = 0;
    complete = false;
}

namespace {

// The launch plan of a single mesh over its current device arrays (drops captured graphs).
hipError_t mesh_plan(md_mesh *m, hipStream_t st) {
    std::vector<const MeshDev *> ms{&m->dev};
    std::vector<const Factor *> fs{&m->fac};
    m->plan.pcg = m->cfg.solver == MD_SOLVER_PCG;
    m->plan.pcg_iters = m->cfg.pcg_max_iters;
    return build_plan(m->plan, ms, fs, m->fp64(), m->cfg.refine_steps, st);
}

// Moves every device pointer of the mesh from the old arena bases to the new ones.
void rebind(md_mesh *m, const char *oa, size_t asz, char *na, const char *ol, size_t lsz, char *nl) {
    auto fix = [&](auto &p) {
        using P = std::remove_reference_t<decltype(p)>;
        const char *c = reinterpret_cast<const char *>(p);
        if (!c) return;
        if (c >= oa && c < oa + asz) p = reinterpret_cast<P>(na + (c - oa));
        else if (c >= ol && c < ol + lsz) p = reinterpret_cast<P>(nl + (c - ol));
    };
    MeshDev &d = m->dev;
    fix(d.slice_ptr); fix(d.sell_col); fix(d.a32); fix(d.a64); fix(d.w64); fix(d.tri); fix(d.fe_val);
    fix(d.slot_fe_ptr); fix(d.slot_fe); fix(d.p); fix(d.pp); fix(d.quat); fix(d.r); fix(d.z);
    fix(d.upart); fix(d.xpart); fix(d.child_ptr); fix(d.child); fix(d.udst); fix(d.rowinfo); fix(d.xloc);
    fix(d.pc); fix(d.pr); fix(d.ftile); fix(d.btile); fix(d.sub); fix(d.sub_h); fix(d.snode); fix(d.scp);
    fix(d.sfc); fix(d.sbo); fix(d.srow); fix(d.sfi); fix(d.lrows); fix(d.energy_part); fix(d.w32);
    fix(d.diag); fix(d.cg_x); fix(d.cg_p); fix(d.cg_v); fix(d.cg_part);
    fix(m->d_fac); fix(m->d_hlist); fix(m->d_offrows); fix(m->d_rel); fix(m->d_tgt_xyz);
    fix(m->d_orig); fix(m->d_xyz);
}

// True when an 8-byte word of the mesh's device record still addresses an old arena (a
// pointer field rebind() does not know): checked whenever the old and new ranges are apart.
bool stale_pointer(const md_mesh *m, const char *oa, size_t asz, const char *ol, size_t lsz) {
    const unsigned char *b = reinterpret_cast<const unsigned char *>(&m->dev);
    for (size_t o = 0; o + sizeof(uintptr_t) <= sizeof(MeshDev); o += sizeof(uintptr_t)) {
        uintptr_t w;
        std::memcpy(&w, b + o, sizeof(w));
        const char *c = reinterpret_cast<const char *>(w);
        if ((c >= oa && c < oa + asz) || (c >= ol && c < ol + lsz)) return true;
    }
    return false;
}

}  // namespace

int mesh_evict(md_mesh *m, hipStream_t st) {
    if (m->evicted) return MD_OK;
    if (!m->analyzed || !m->arena.allocated()) { set_error("mesh_evict: the mesh has no device state"); return MD_ERR_STATE; }
    HostImage &I = m->img;
    const size_t asz = m->arena.size(), lsz = m->larena.size();
    if (I.a && (I.asz != asz || I.lsz != lsz)) I.release();
    if (!I.a) {
        if (hipHostMalloc(&I.a, std::max<size_t>(asz, 1), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&I.l, std::max<size_t>(lsz, 1), hipHostMallocDefault) != hipSuccess) {
            I.release();
            set_error("mesh_evict: pinned host image of " + std::to_string(asz + lsz) + " bytes: allocation failed");
            return MD_ERR_NOMEM;
        }
        I.asz = asz;
        I.lsz = lsz;
        I.complete = false;
    }
    char *base = m->arena.base();
    hipError_t e = hipSuccess;
    if (!I.complete) {
        e = hipMemcpyAsync(I.a, base, asz, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && lsz) e = hipMemcpyAsync(I.l, m->larena.base(), lsz, hipMemcpyDeviceToHost, st);
    } else {
        // the image holds every static array already: only the iteration state moved since
        const size_t vsz = m->fp64() ? sizeof(d4) : sizeof(f4);
        const size_t opp = reinterpret_cast<char *>(m->dev.pp) - base;
        const size_t oq = reinterpret_cast<char *>(m->dev.quat) - base;
        e = hipMemcpyAsync(static_cast<char *>(I.a) + opp, m->dev.pp, sizeof(d4) * m->n, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(static_cast<char *>(I.a) + oq, m->dev.quat, vsz * m->n, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { set_error(std::string("mesh_evict: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    I.complete = true;
    m->img_old_a = base;
    m->img_old_l = m->larena.base();
    m->plan.drop_graph();
    m->arena.release();
    m->larena.release();
    m->release_fac_ws();
    m->evicted = true;
    ++m->gen;
    return MD_OK;
}

int mesh_restore(md_mesh *m, hipStream_t st) {
    if (!m->evicted) return MD_OK;
    HostImage &I = m->img;
    if (m->arena.allocate() != hipSuccess || m->larena.allocate() != hipSuccess) {
        m->arena.release();
        m->larena.release();
        set_error("mesh_restore: device allocation of " + std::to_string(I.asz + I.lsz) + " bytes failed");
        return MD_ERR_NOMEM;
    }
    char *na = m->arena.base(), *nl = m->larena.base();
    hipError_t e = hipMemcpyAsync(na, I.a, I.asz, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && I.lsz) e = hipMemcpyAsync(nl, I.l, I.lsz, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        m->arena.release();
        m->larena.release();
        set_error(std::string("mesh_restore: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    const char *oa = m->img_old_a, *ol = m->img_old_l;
    rebind(m, oa, I.asz, na, ol, I.lsz, nl);
    const bool apart = (na + I.asz <= oa || oa + I.asz <= na) && (!I.lsz || nl + I.lsz <= ol || ol + I.lsz <= nl);
    if (apart && stale_pointer(m, oa, I.asz, ol, I.lsz)) {
        m->arena.release();
        m->larena.release();
        set_error("mesh_restore: a device pointer of the mesh was not re-pointed");
        return MD_ERR_STATE;
    }
    m->evicted = false;
    ++m->gen;
    e = mesh_plan(m, st);
    if (e != hipSuccess) { set_error(std::string("mesh_restore: launch plan: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    return MD_OK;
}

int ensure_resident(md_mesh *m, const char *who) {
    if (m->servers > 0) {
        set_error(std::string(who) + ": the mesh is registered with a server (its results come from "
                  "md_server_wait / md_server_poll until the server is destroyed)");
        return MD_ERR_STATE;
    }
    if (!m->evicted) return MD_OK;
    int rc = mesh_restore(m, m->ctx->stream);
    if (rc) return rc;
    m->img.release();  // a mesh outside a server may be re-prepared: the image would go stale
    return MD_OK;
}

}  // namespace md

namespace {

// Checks the candidate budgets of an autotune call.
int check_budgets(const char *who, const int *budgets, int count, int iterations) {
    if (!budgets || count <= 0 || iterations <= 0) {
        set_error(std::string(who) + ": need at least one candidate and one iteration");
        return MD_ERR_INVALID;
    }
    for (int i = 0; i < count; ++i)
        if (budgets[i] < 0 || budgets[i] > kSubBudgetMax) {
            set_error(std::string(who) + ": budget " + std::to_string(budgets[i]) + " outside [0, " +
                      std::to_string(kSubBudgetMax) + "]");
            return MD_ERR_INVALID;
        }
    return MD_OK;
}

}  // namespace

extern "C" int md_mesh_autotune(md_mesh *m, const int *budgets, int count, int iterations, int *best_budget,
                                double *best_ms) {
    if (!m) { set_error("md_mesh_autotune: null mesh"); return MD_ERR_INVALID; }
    int rc = check_budgets("md_mesh_autotune", budgets, count, iterations);
    if (rc) return rc;
    if (m->dirty || !m->analyzed || !m->factored) { set_error("md_mesh_autotune: mesh is not prepared"); return MD_ERR_STATE; }
    if (m->cfg.solver != MD_SOLVER_LDLT) { set_error("md_mesh_autotune: only the LDLT solver has a sweep layout"); return MD_ERR_INVALID; }
    if (m->servers > 0) { set_error("md_mesh_autotune: the mesh is registered with a server"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if ((rc = ensure_resident(m, "md_mesh_autotune"))) return rc;
    hipStream_t st = m->ctx->stream;
    std::vector<md_mesh *> ms{m};
    StateSnapshot snap;
    MD_HIP(snapshot_state(ms, st, snap));
    const bool g = m->cfg.use_graph != 0;
    int best = m->cfg.subtree_budget;
    double best_t = 1e300;
    for (int i = 0; i < count; ++i) {
        rc = relayout(m, budgets[i]);
        if (rc) return rc;
        double t = 0.0;
        MD_HIP(run_iterations(m->plan, st, 1, g));  // first replay captures the graphs
        MD_HIP(restore_state(ms, st, snap));
        MD_HIP(time_plan(m->plan, st, iterations, g, t));
        MD_HIP(restore_state(ms, st, snap));
        if (t < best_t) { best_t = t; best = budgets[i]; }
    }
    rc = relayout(m, best);
    if (rc) return rc;
    if (best_budget) *best_budget = best;
    if (best_ms) *best_ms = best_t / iterations;
    return MD_OK;
}

extern "C" int md_mesh_autotune_factor(md_mesh *m, const double *ratios, int nratios, const int *budgets,
                                       int nbudgets, int iterations, double *best_ratio, int *best_budget,
                                       double *best_ms) {
    if (!m || !ratios || nratios <= 0) { set_error("md_mesh_autotune_factor: need at least one ratio"); return MD_ERR_INVALID; }
    int rc = check_budgets("md_mesh_autotune_factor", budgets, nbudgets, iterations);
    if (rc) return rc;
    for (int i = 0; i < nratios; ++i)
        if (!(ratios[i] >= 0.0 && ratios[i] <= 1.0)) {
            set_error("md_mesh_autotune_factor: ratios must lie in [0, 1]");
            return MD_ERR_INVALID;
        }
    if (m->dirty || !m->analyzed || !m->factored || m->iterated) {
        set_error("md_mesh_autotune_factor: call it on a freshly prepared mesh (md_mesh_prepare, no iteration yet)");
        return MD_ERR_STATE;
    }
    if (m->cfg.solver != MD_SOLVER_LDLT) { set_error("md_mesh_autotune_factor: only the LDLT solver has a factor layout"); return MD_ERR_INVALID; }
    if (m->servers > 0) { set_error("md_mesh_autotune_factor: the mesh is registered with a server"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if ((rc = ensure_resident(m, "md_mesh_autotune_factor"))) return rc;
    // every candidate re-runs the dirty setup from the rest pose: the state it leaves is the
    // freshly prepared one, whatever the candidate
    const std::vector<double> rest = m->rest;
    auto reprepare = [&](double ratio) -> int {
        m->cfg.amalg_ratio = ratio;
        m->dirty = true;
        m->con_set_changed = true;
        m->initialized = false;
        bool early = false;
        int r = setup(m, rest.data(), &early);
        if (r == MD_OK && early) { set_error("md_mesh_autotune_factor: no constrained vertex"); r = MD_ERR_STATE; }
        if (r == MD_OK) ++m->gen;
        return r;
    };
    double bt = 1e300, br = m->cfg.amalg_ratio;
    int bb = m->cfg.subtree_budget;
    for (int i = 0; i < nratios; ++i) {
        if ((rc = reprepare(ratios[i]))) return rc;
        int b = 0;
        double t = 0.0;
        if ((rc = md_mesh_autotune(m, budgets, nbudgets, iterations, &b, &t))) return rc;
        if (t < bt) { bt = t; br = ratios[i]; bb = b; }
    }
    m->cfg.subtree_budget = bb;
    if ((rc = reprepare(br))) return rc;
    if (best_ratio) *best_ratio = br;
    if (best_budget) *best_budget = bb;
    if (best_ms) *best_ms = bt;
    return MD_OK;
}

extern "C" int md_mesh_prepare(md_mesh *m, const double *xyz) {
    if (!m || !xyz) { set_error("md_mesh_prepare: null pointer"); return MD_ERR_INVALID; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_prepare")) return rc;
    if (!m->dirty) return MD_OK;
    bool early = false;
    int rc = setup(m, xyz, &early);
    if (rc) return rc;
    if (early) { set_error("md_mesh_prepare: no constrained vertex"); return MD_ERR_STATE; }
    return MD_OK;
}

extern "C" int md_meshes_prepare(md_ctx *ctx, int count, md_mesh *const *meshes, const double *const *xyz,
                                 int nthreads) {
    if (!ctx || count < 0 || (count > 0 && (!meshes || !xyz))) {
        set_error("md_meshes_prepare: invalid argument");
        return MD_ERR_INVALID;
    }
    for (int k = 0; k < count; ++k)
        if (!meshes[k] || meshes[k]->ctx != ctx || !xyz[k]) {
            set_error("md_meshes_prepare: mesh " + std::to_string(k) + " is null or from another context");
            return MD_ERR_INVALID;
        }
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = std::min(nthreads, 16);
    hipSetDevice(ctx->device);
    for (int k = 0; k < count; ++k)
        if (int rc = ensure_resident(meshes[k], "md_meshes_prepare")) return rc;
    std::vector<md_mesh *> todo;
    std::vector<Prep> preps;
    for (int k = 0; k < count; ++k)
        if (meshes[k]->dirty) {
            todo.push_back(meshes[k]);
            preps.emplace_back();
            prep_begin(meshes[k], xyz[k], preps.back());
        }
    int nt = (int)todo.size();
    parallel_for(nt, nthreads, [&](int k) {
        if (preps[k].need_analysis) {
            auto t0 = Clock::now();
            analyze_host(todo[k]);
            todo[k]->setup_ms_host = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        }
    });
    for (int k = 0; k < nt; ++k) {
        int rc = prep_device_a(todo[k], preps[k]);
        if (rc) return rc;
        if (preps[k].early) {
            set_error("md_meshes_prepare: mesh without constrained vertices");
            return MD_ERR_STATE;
        }
    }
    std::vector<std::string> errs(nt);
    std::vector<char> bad(nt, 0);
    parallel_for(nt, nthreads, [&](int k) {
        if (preps[k].need_factor && todo[k]->cfg.solver == MD_SOLVER_LDLT && !todo[k]->cfg.factor_device)
            bad[k] = !factor_host(todo[k], errs[k]);
    });
    for (int k = 0; k < nt; ++k)
        if (bad[k]) { set_error(errs[k]); return MD_ERR_FACTOR; }
    for (int k = 0; k < nt; ++k) {
        int rc = prep_device_b(todo[k], preps[k]);
        if (rc) return rc;
    }
    return MD_OK;
}

extern "C" int md_mesh_deform(md_mesh *m, int iterations, double *xyz, int *ok) {
    if (ok) *ok = 0;
    if (!m || !xyz || iterations < 0) { set_error("md_mesh_deform: invalid argument"); return MD_ERR_INVALID; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_deform")) return rc;
    auto ms_since = [](Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); };
    const auto t_call = Clock::now();
    for (double &x : m->deform_ms) x = 0.0;
    if (m->dirty) {
        bool early = false;
        int rc = setup(m, xyz, &early);
        for (int k = 0; k < 3; ++k) m->deform_ms[k] = m->setup_phase_ms[k];
        if (rc == MD_ERR_FACTOR) return rc;   // deform() == false (arap.h:116-117)
        if (rc) return rc;
        if (early) { if (ok) *ok = 1; m->deform_ms[5] = ms_since(t_call); return MD_OK; }
    }
    auto t0 = Clock::now();
    int rc = iterate(m, iterations);  // arap.h:122-129
    if (rc) return rc;
    m->deform_ms[3] = ms_since(t0);
    t0 = Clock::now();
    hipError_t e = download_positions(m, xyz);  // write-back, arap.h:133-135
    if (e != hipSuccess) { set_error("position download failed"); return MD_ERR_HIP; }
    m->deform_ms[4] = ms_since(t0);
    m->deform_ms[5] = ms_since(t_call);
    if (ok) *ok = 1;
    return MD_OK;
}

extern "C" int md_mesh_iterate(md_mesh *m, int iterations) {
    if (!m || iterations < 0) { set_error("md_mesh_iterate: invalid argument"); return MD_ERR_INVALID; }
    if (m->dirty) { set_error("md_mesh_iterate: mesh is not prepared"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_iterate")) return rc;
    return iterate(m, iterations);
}

extern "C" int md_mesh_download(md_mesh *m, double *xyz) {
    if (!m || !xyz) { set_error("md_mesh_download: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_download: no device state yet"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_download")) return rc;
    if (download_positions(m, xyz) != hipSuccess) { set_error("position download failed"); return MD_ERR_HIP; }
    return MD_OK;
}

extern "C" int md_mesh_cotan_nnz(md_mesh *m, int64_t *nnz) {
    if (!m || !nnz) { set_error("md_mesh_cotan_nnz: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_cotan_nnz: call deform first"); return MD_ERR_STATE; }
    int64_t c = 0;
    const Topology &t = m->topo;
    for (int s = 0; s < t.nslices; ++s)
        for (int l = 0; l < kSlice; ++l) {
            int iv = s * kSlice + l;
            if (iv >= t.n) break;
            for (int q = t.slice_ptr[s]; q < t.slice_ptr[s + 1]; ++q)
                if (t.sell_col[(int64_t)q * kSlice + l] != iv) ++c;
        }
    *nnz = c;
    return MD_OK;
}

extern "C" int md_mesh_cotan_csr(md_mesh *m, int *rowptr, int *col, double *val) {
    if (!m || !rowptr || !col || !val) { set_error("md_mesh_cotan_csr: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_cotan_csr: call deform first"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_cotan_csr")) return rc;
    const Topology &t = m->topo;
    int64_t slots = t.total_slots();
    std::vector<double> w(slots);
    if (copy_on_stream(m, w.data(), m->dev.w64, sizeof(double) * slots, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("download of cotan weights failed");
        return MD_ERR_HIP;
    }
    rowptr[0] = 0;
    int64_t nnz = 0;
    for (int v = 0; v < m->n; ++v) {
        int iv = t.internal_of[v];
        int s = iv / kSlice, l = iv % kSlice;
        for (int q = t.slice_ptr[s]; q < t.slice_ptr[s + 1]; ++q) {
            int64_t g = (int64_t)q * kSlice + l;
            int j = t.sell_col[g];
            if (j == iv) continue;
            col[nnz] = t.original_of[j];
            val[nnz] = w[g];
            ++nnz;
        }
        rowptr[v + 1] = (int)nnz;
    }
    return MD_OK;
}

extern "C" int md_mesh_rotations(md_mesh *m, double *R) {
    if (!m || !R) { set_error("md_mesh_rotations: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_rotations: call deform first"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_rotations")) return rc;
    int n = m->n;
    std::vector<double> q(4 * (size_t)n);
    if (m->fp64()) {
        if (copy_on_stream(m, q.data(), m->dev.quat, sizeof(d4) * n, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("rotation download failed"); return MD_ERR_HIP;
        }
    } else {
        std::vector<float> qf(4 * (size_t)n);
        if (copy_on_stream(m, qf.data(), m->dev.quat, sizeof(f4) * n, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("rotation download failed"); return MD_ERR_HIP;
        }
        for (size_t k = 0; k < qf.size(); ++k) q[k] = qf[k];
    }
    for (int iv = 0; iv < n; ++iv) {
        double x = q[4 * iv], y = q[4 * iv + 1], z = q[4 * iv + 2], w = q[4 * iv + 3];
        double s = 2.0 / (x * x + y * y + z * z + w * w);
        double *o = R + 9 * (size_t)m->topo.original_of[iv];
        o[0] = 1 - s * (y * y + z * z); o[1] = s * (x * y - w * z); o[2] = s * (x * z + w * y);
        o[3] = s * (x * y + w * z); o[4] = 1 - s * (x * x + z * z); o[5] = s * (y * z - w * x);
        o[6] = s * (x * z - w * y); o[7] = s * (y * z + w * x); o[8] = 1 - s * (x * x + y * y);
    }
    return MD_OK;
}

extern "C" int md_mesh_normals(md_mesh *m, double *nrm) {
    if (!m || !nrm) { set_error("md_mesh_normals: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_normals: call deform first"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_normals")) return rc;
    hipStream_t st = m->ctx->stream;
    double *fn = nullptr;
    d4 *out = nullptr;
    MD_HIP(hipMalloc((void **)&fn, sizeof(double) * 3 * std::max(1, m->F)));
    hipError_t e = hipMalloc((void **)&out, sizeof(d4) * std::max(1, m->n));
    if (e == hipSuccess) {
        const Plan &P = m->plan;
        launch_normals(st, P.set(P.off_face), P.blk_face, P.set(P.off_slice), P.blk_slice, fn, out);
        e = hipGetLastError();
    }
    std::vector<d4> h(m->n);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), out, sizeof(d4) * m->n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hipStreamSynchronize(st);
    if (fn) hipFree(fn);
    if (out) hipFree(out);
    if (e != hipSuccess) { set_error(std::string("md_mesh_normals: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    for (int iv = 0; iv < m->n; ++iv) {
        double *o = nrm + 3 * (size_t)m->topo.original_of[iv];
        o[0] = h[iv].x; o[1] = h[iv].y; o[2] = h[iv].z;
    }
    return MD_OK;
}

namespace {
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
// Workspace of md_mesh_normals_noalloc: face normals, vertex normals (internal order), the
// internal-to-original permutation and the packed result.
size_t normals_scratch_bytes(size_t n, size_t F) {
    return align256(sizeof(double) * 3 * F) + align256(sizeof(d4) * n) + align256(sizeof(int) * n) +
           align256(sizeof(double) * 3 * n);
}
}  // namespace

extern "C" int md_mesh_normals_noalloc(md_mesh *m, double *nrm) {
    if (!m || !nrm) { set_error("md_mesh_normals_noalloc: null pointer"); return MD_ERR_INVALID; }
    const size_t need = normals_scratch_bytes((size_t)m->n, (size_t)m->F);
    if (need > kScratchBytes) {
        set_error("md_mesh_normals_noalloc: a mesh of " + std::to_string(m->n) + " vertices and " +
                  std::to_string(m->F) + " faces needs " + std::to_string(need) +
                  " bytes of device workspace, more than its maximum of " + std::to_string(kScratchBytes) +
                  " (md_mesh_normals has no limit)");
        return MD_ERR_INVALID;
    }
    if (!m->analyzed) { set_error("md_mesh_normals_noalloc: call deform first"); return MD_ERR_STATE; }
    if (m->servers > 0) {
        set_error("md_mesh_normals_noalloc: the mesh is registered with a server");
        return MD_ERR_STATE;
    }
    if (m->evicted) {
        set_error("md_mesh_normals_noalloc: the mesh is evicted to host memory (restoring it allocates)");
        return MD_ERR_STATE;
    }
    hipSetDevice(m->ctx->device);
    hipStream_t st = m->ctx->stream;
    char *base = static_cast<char *>(m->ctx->scratch);
    const size_t n = (size_t)m->n, F = (size_t)m->F;
    double *fn = reinterpret_cast<double *>(base);
    d4 *vn = reinterpret_cast<d4 *>(base + align256(sizeof(double) * 3 * F));
    int *perm = reinterpret_cast<int *>(reinterpret_cast<char *>(vn) + align256(sizeof(d4) * n));
    double *dst = reinterpret_cast<double *>(reinterpret_cast<char *>(perm) + align256(sizeof(int) * n));
    const Plan &P = m->plan;
    launch_normals(st, P.set(P.off_face), P.blk_face, P.set(P.off_slice), P.blk_slice, fn, vn);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && n)
        e = hipMemcpyAsync(perm, m->topo.original_of.data(), sizeof(int) * n, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        launch_pack_xyz(st, vn, perm, m->n, dst);
        e = hipGetLastError();
    }
    if (e == hipSuccess && n) e = hipMemcpyAsync(nrm, dst, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { set_error(std::string("md_mesh_normals_noalloc: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    return MD_OK;
}

extern "C" int md_mesh_energy(md_mesh *m, double *energy) {
    if (!m || !energy) { set_error("md_mesh_energy: null pointer"); return MD_ERR_INVALID; }
    if (!m->analyzed) { set_error("md_mesh_energy: call deform first"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_energy")) return rc;
    const Plan &P = m->plan;
    launch_energy(m->ctx->stream, P.set(P.off_slice), P.blk_slice, m->fp64());
    int nb = std::max(1, cdiv(m->topo.nslices, kSlicesPerBlock));
    std::vector<double> part(nb);
    MD_HIP(hipMemcpyAsync(part.data(), m->dev.energy_part, sizeof(double) * nb, hipMemcpyDeviceToHost,
                          m->ctx->stream));
    MD_HIP(hipStreamSynchronize(m->ctx->stream));
    double e = 0.0;
    for (int b = 0; b < P.blk_slice; ++b) e += part[b];
    *energy = e;
    return MD_OK;
}

extern "C" int md_mesh_get_stats(md_mesh *m, md_mesh_stats *out) {
    if (!m || !out) { set_error("md_mesh_get_stats: null pointer"); return MD_ERR_INVALID; }
    std::memset(out, 0, sizeof(*out));
    out->n = m->n;
    out->F = m->F;
    out->subtree_budget = m->cfg.subtree_budget;
    if (!m->analyzed) return MD_OK;
    out->subtrees = m->fac.nsub;
    AnalysisStats s = stats(m->topo, m->fac);
    out->nf = s.nf;
    out->nslices = s.nslices;
    out->supernodes = s.ns;
    out->levels = s.nlevels;
    out->max_sn_width = s.max_width;
    out->max_sn_rows = s.max_rows;
    out->slots = s.slots;
    int64_t nnz = 0;
    md_mesh_cotan_nnz(m, &nnz);
    out->cotan_nnz = nnz;
    out->panel_entries = s.panel_entries;
    out->factor_nnz = s.strict_nnz_equiv;
    out->factor_nnz_exact = s.nnz_exact;
    out->factor_ms = m->factor_ms;
    for (int k = 0; k < 6; ++k) out->deform_ms[k] = m->deform_ms[k];
    out->fwd_items = s.fwd_items;
    out->bwd_items = s.bwd_items;
    out->device_bytes = (int64_t)(m->arena.size() + m->larena.size() + m->fac_ws_bytes);
    out->setup_ms_host = m->setup_ms_host;
    // algorithmic bytes per iteration: each array element a phase must touch, counted once
    const Topology &t = m->topo;
    const Factor &f = m->fac;
    const int64_t es = m->fp64() ? 8 : 4;      // bytes per panel / edge scalar
    const int64_t vs = 4 * es;                 // bytes per solve vector entry (x, y, z, pad)
    const int64_t ps = 32;                     // FP64 position (x, y, z, pad)
    int64_t slots = t.total_slots();
    int64_t free_slots = t.nslices > 0 ? (int64_t)t.slice_ptr[(t.nf + kSlice - 1) / kSlice] * kSlice : 0;
    // local: neighbour ids + weighted rest edges per slot, p' and the rotation per vertex
    out->bytes_local = slots * (4 + 3 * es) + (int64_t)t.n * (ps + vs);
    // residual: ids, edges and FP64 weights per slot, p' and rotation per vertex, residual out
    out->bytes_residual = free_slots * (4 + 3 * es + 8) + (int64_t)t.n * (ps + vs) + (int64_t)t.nf * vs;
    int64_t gw = 0, dw = 0, upd = 0, xp = 0, offr = 0, xloc_reads = 0, lrows = 0;
    for (int k = 0; k < f.ns; ++k) {
        gw += (int64_t)f.noff[k] * f.w[k];
        dw += (int64_t)f.w[k] * f.w[k];
        upd += (int64_t)f.kc[k] * f.noff[k];
        offr += f.noff[k];
        if (f.level[k] >= 0) {  // subtree rows' x stays in LDS and goes straight into p'
            xp += (int64_t)f.kr[k] * f.w[k];
            lrows += f.w[k];
        }
    }
    // solution partials gathered from HBM: off rows of level-swept supernodes, and off rows
    // of subtree supernodes that lie outside their subtree (the others are read from LDS)
    for (int k = 0; k < f.ns; ++k)
        if (f.level[k] >= 0)
            for (int e = f.roff[k]; e < f.roff[k + 1]; ++e) xloc_reads += f.xloc[4 * (size_t)e + 2];
    for (size_t i = 0; i < f.sbo.size(); i += 2)
        if (f.sbo[i + 1] >= 0) xloc_reads += f.sbo[i + 1] >> 16;
    // index data of the bottom-subtree sweeps: records (both), contribution lists (forward),
    // off-row sources (backward)
    const int64_t sub_fwd_idx = 4 * (int64_t)(f.snode.size() + f.scp.size() + f.sfc.size() + f.sub_h.size());  // scp: 2 ints per position
    const int64_t sub_bwd_idx = 4 * (int64_t)(f.snode.size() + f.sbo.size() + f.sub_h.size());
    // forward: G panel, r in, z out, update partials out (with their slot addresses) and every
    // update slot of the level fronts read back (unused slots included)
    const int64_t uslots = f.ubase.empty() ? 0 : f.ubase.back();
    int64_t udst_reads = 0;
    for (size_t e = 0; e < f.udst.size(); ++e) udst_reads += f.udst[e] >= 0;
    out->bytes_forward = gw * es + 2 * (int64_t)t.nf * vs + upd * vs + uslots * vs + 8 * udst_reads + sub_fwd_idx;
    // backward + update: Dinv and G panels, z in, x partials of ancestors gathered (with
    // their locators), level rows' x partials out and summed (row list, row locators), the
    // p' update
    out->bytes_backward = (dw + gw) * es + (int64_t)t.nf * vs + xloc_reads * vs + offr * 16 + 2 * xp * vs +
                          lrows * (4 + 16) + (int64_t)t.nf * 2 * ps + sub_bwd_idx;
    return MD_OK;
}

extern "C" int md_mesh_time_iterations(md_mesh *m, int warmup, int iterations, double *total_ms,
                                       double *phase_ms4) {
    if (!m || iterations < 0 || warmup < 0) { set_error("md_mesh_time_iterations: invalid argument"); return MD_ERR_INVALID; }
    if (m->dirty) { set_error("md_mesh_time_iterations: mesh is not prepared"); return MD_ERR_STATE; }
    hipSetDevice(m->ctx->device);
    if (int rc = ensure_resident(m, "md_mesh_time_iterations")) return rc;
    hipStream_t st = m->ctx->stream;
    bool g = m->cfg.use_graph != 0;
    MD_HIP(run_iterations(m->plan, st, warmup, g));
    MD_HIP(hipStreamSynchronize(st));
    hipEvent_t a, b;
    MD_HIP(hipEventCreate(&a));
    MD_HIP(hipEventCreate(&b));
    MD_HIP(hipEventRecord(a, st));
    MD_HIP(run_iterations(m->plan, st, iterations, g));
    MD_HIP(hipEventRecord(b, st));
    MD_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    hipEventElapsedTime(&ms, a, b);
    if (total_ms) *total_ms = ms;
    if (phase_ms4) {
        // phases from a separate pass with events between them (no graph)
        hipEvent_t ev[5];
        for (auto &x : ev) hipEventCreate(&x);
        double acc[4] = {0, 0, 0, 0};
        for (int i = 0; i < iterations; ++i) {
            enqueue_iteration(m->plan, st, ev);
            hipEventSynchronize(ev[4]);
            for (int k = 0; k < 4; ++k) {
                float t = 0.f;
                hipEventElapsedTime(&t, ev[k], ev[k + 1]);
                acc[k] += t;
            }
        }
        for (int k = 0; k < 4; ++k) phase_ms4[k] = acc[k];
        for (auto &x : ev) hipEventDestroy(x);
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    m->iterated = true;
    return MD_OK;
}

extern "C" int md_copy_probe(md_ctx *ctx, size_t bytes, int repeats, double *gbps) {
    if (!ctx || bytes < 16 || repeats < 1) { set_error("md_copy_probe: invalid argument"); return MD_ERR_INVALID; }
    hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    const long long n16 = (long long)(bytes / 16);
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, (size_t)n16 * 16);
    if (e == hipSuccess) e = hipMalloc(&b, (size_t)n16 * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, (size_t)n16 * 16, st);
    if (e == hipSuccess) launch_copy_probe(st, a, b, n16);  // warm-up
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int i = 0; e == hipSuccess && i < repeats; ++i) launch_copy_probe(st, a, b, n16);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    if (e != hipSuccess) { set_error(std::string("md_copy_probe: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    if (gbps) *gbps = 2.0 * 16.0 * (double)n16 * repeats / (ms * 1e-3) / 1e9;
    return MD_OK;
}

