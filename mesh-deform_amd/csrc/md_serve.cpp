// Batches: many prepared meshes iterated by the same launches (one launch per kernel per
// level covers every mesh of the batch), replayed from one captured hipGraph.
#include <algorithm>
#include <memory>

#include "md_runtime.h"

using namespace md;

struct md_batch {
    md_ctx *ctx = nullptr;
    std::vector<md_mesh *> meshes;
    Plan plan;
    bool use_graph = true;
};

extern "C" int md_batch_create(md_ctx *ctx, int count, md_mesh *const *meshes, md_batch **out) {
    if (!ctx || !out || count <= 0 || !meshes) { set_error("md_batch_create: invalid argument"); return MD_ERR_INVALID; }
    *out = nullptr;
    std::unique_ptr<md_batch> b(new md_batch());
    b->ctx = ctx;
    std::vector<const MeshDev *> devs;
    std::vector<const Factor *> facs;
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
        devs.push_back(&m->dev);
        facs.push_back(&m->fac);
    }
    b->use_graph = meshes[0]->cfg.use_graph != 0;
    hipSetDevice(ctx->device);
    hipError_t e = build_plan(b->plan, devs, facs, prec == MD_PREC_FP64, refine, ctx->stream);
    if (e != hipSuccess) { set_error(std::string("md_batch_create: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    *out = b.release();
    return MD_OK;
}

extern "C" int md_batch_iterate(md_batch *b, int iterations) {
    if (!b || iterations < 0) { set_error("md_batch_iterate: invalid argument"); return MD_ERR_INVALID; }
    hipSetDevice(b->ctx->device);
    hipError_t e = run_iterations(b->plan, b->ctx->stream, iterations, b->use_graph);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) { set_error(std::string("md_batch_iterate: ") + hipGetErrorString(e)); return MD_ERR_HIP; }
    for (md_mesh *m : b->meshes) m->iterated = true;
    return MD_OK;
}

extern "C" int md_batch_time_iterations(md_batch *b, int warmup, int iterations, double *total_ms) {
    if (!b || iterations < 0 || warmup < 0) { set_error("md_batch_time_iterations: invalid argument"); return MD_ERR_INVALID; }
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

extern "C" int md_batch_destroy(md_batch *b) {
    if (!b) return MD_OK;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    delete b;
    return MD_OK;
}

// ---- server (continuous batching) ---------------------------------------------------------

struct md_server {
    md_ctx *ctx = nullptr;
};

extern "C" int md_server_create(md_ctx *ctx, const md_server_config *, md_server **out) {
    (void)ctx; (void)out;
    set_error("md_server_create: not available yet");
    return MD_ERR_STATE;
}
extern "C" int md_server_add_mesh(md_server *, md_mesh *, int *) { set_error("server not available"); return MD_ERR_STATE; }
extern "C" int md_server_submit(md_server *, int, int, const int *, const double *, int, uint64_t *) {
    set_error("server not available"); return MD_ERR_STATE;
}
extern "C" int md_server_poll(md_server *, uint64_t, int *, double *, double *) { set_error("server not available"); return MD_ERR_STATE; }
extern "C" int md_server_wait(md_server *, uint64_t, double, double *, double *) { set_error("server not available"); return MD_ERR_STATE; }
extern "C" int md_server_destroy(md_server *) { return MD_OK; }
