/*
 * md_arap.h -- C ABI of libmdarap.so, the MI355X (gfx950) As-Rigid-As-Possible
 * deformation library.
 *
 * The ABI is the drop-in boundary for the deformation path of cheind/mesh-deform:
 * include/deform/arap.h (the C++ class AsRigidAsPossibleDeformation) forwards every call
 * to it, and the Python tests and bench drive it through ctypes. Signatures use plain C
 * types only: pointers to host memory, sizes and opaque handles. Every function returns an
 * md_status (0 = success); md_last_error() gives a message for the calling thread.
 *
 * Threading: a context and everything created from it belong to one host thread at a time
 * (the reference class is single-threaded too). A server (md_server_*) owns an internal
 * worker thread and a stream of its own; its submit/poll/wait/stats calls may come from any
 * thread. A mesh registered with a server belongs to it: md_mesh_* calls that change or read
 * the mesh's device state, md_mesh_destroy and batches holding the mesh return MD_ERR_STATE
 * until the server is destroyed.
 * Ownership: the caller owns every host buffer it passes; the library owns device memory.
 */
#ifndef MD_ARAP_H
#define MD_ARAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MD_OK = 0,
    MD_ERR_INVALID = 1,   /* bad argument: null pointer, size, out-of-range vertex index */
    MD_ERR_NO_DEVICE = 2, /* no usable GPU, or the device index is out of range */
    MD_ERR_HIP = 3,       /* a HIP runtime call failed */
    MD_ERR_FACTOR = 4,    /* the system is not positive definite (deform() returns false) */
    MD_ERR_NOMEM = 5,     /* host or device allocation failed */
    MD_ERR_BUDGET = 6,    /* the request does not fit the server's HBM budget */
    MD_ERR_STATE = 7,     /* the call is not valid in the object's current state */
    MD_ERR_TIMEOUT = 8    /* md_server_wait timed out */
} md_status;

typedef enum {
    MD_PREC_FP64 = 0,        /* every step in FP64 (the reference's double PrecisionType) */
    MD_PREC_FP32_REFINE = 1  /* FP32 local step and solve, FP64 factor and residual */
} md_precision;

typedef enum {
    MD_SOLVER_LDLT = 0, /* supernodal block LDL^T, level-scheduled triangular sweeps */
    MD_SOLVER_PCG = 1   /* Jacobi-preconditioned CG on the residual, warm-started */
} md_solver;

typedef enum {
    MD_REST_REFERENCE = 0, /* each dirty deform() takes the current mesh as the rest pose
                              (inc/deform/arap.h:102-120,162-168) */
    MD_REST_FIXED = 1      /* the rest pose, weights and factor of the first deform() are
                              kept; later calls only move targets and warm-start */
} md_rest_mode;

typedef struct {
    int precision;     /* md_precision, default MD_PREC_FP64 */
    int solver;        /* md_solver, default MD_SOLVER_LDLT */
    int refine_steps;  /* residual/solve steps per ARAP iteration (>= 1), default 1 */
    int rest_mode;     /* md_rest_mode, default MD_REST_REFERENCE */
    int pcg_max_iters; /* MD_SOLVER_PCG: CG iterations per solve, default 200 */
    double pcg_tol;    /* MD_SOLVER_PCG: relative residual target, default 1e-7 */
    int use_graph;     /* replay iterations from captured hipGraphs (one of 8 iterations, one of a
                          single iteration for the rest; PCG: single only), captured at the
                          first md_mesh_iterate / md_batch_iterate after a (re)layout; default 1 */
    int subtree_budget;/* MD_SOLVER_LDLT: LDS vector entries (at most 1024) of one bottom
                          subtree of the factor swept by a single wavefront (fronts + columns of
                          its supernodes); 0 sweeps every supernode by level launches;
                          default 512 */
    double amalg_ratio;/* MD_SOLVER_LDLT: relaxed supernode amalgamation, the share of explicit
                          zeros a merge may leave (half of it for merges wider than 48
                          columns), in [0, 1]; larger gives fewer, wider supernodes and fewer
                          sweep levels for more panel bytes; default 0.3 */
    int factor_device; /* MD_SOLVER_LDLT: 1 (default) runs the FP64 numeric factorization on the
                          device (multifrontal, one height of the supernodal tree per launch
                          pair); 0 runs it on the host */
    int level_pairs;   /* MD_SOLVER_LDLT: 1 orders the free vertices so that every other sweep level
                          merges into its parents (about half the level launches, denser panels,
                          the same fill); 0 keeps the plain nested-dissection order; default 1 */
} md_config;

typedef struct {
    int n, nf, F;            /* vertices, free vertices, faces */
    int nslices;             /* 64-vertex one-ring slices */
    int supernodes, levels;  /* supernodal tree size and height */
    int max_sn_width, max_sn_rows;
    int64_t slots;           /* one-ring slots incl. padding (64 * slice widths) */
    int64_t cotan_nnz;       /* nonzeros of the cotan weight matrix (2E) */
    int64_t panel_entries;   /* stored factor entries (Dinv + G blocks) */
    int64_t factor_nnz;      /* strictly-lower entries of the supernodal factor as stored
                                (explicit zeros of relaxed amalgamation included) */
    int64_t fwd_items, bwd_items;
    int64_t device_bytes;    /* device memory held by the mesh (with the FP64 factorization
                                workspace that a reference-rest-mode mesh keeps between calls) */
    double setup_ms_host;    /* last analysis + factorization time on the host */
    /* algorithmic bytes of one ARAP iteration with one refinement step, per phase: every
     * array element the phase must read or write counted once */
    int64_t bytes_local, bytes_residual, bytes_forward, bytes_backward;
    int subtree_budget;      /* sweep layout in use (md_config.subtree_budget or the autotuned one) */
    int subtrees;            /* bottom subtrees swept by one wavefront each */
    int64_t factor_nnz_exact;/* strictly-lower nonzeros of the exact factor (no explicit zeros) */
    double factor_ms;        /* wall time of the last numeric factorization (device or host) */
    /* wall milliseconds of the last md_mesh_deform call by phase: [0] host analysis (a new
     * constrained set only), [1] rest pose and initial guess upload with the cotan weights,
     * [2] numeric factorization, [3] iterations, [4] position download and write-back,
     * [5] the whole call */
    double deform_ms[6];
} md_mesh_stats;

/* ---- library ------------------------------------------------------------------------ */
void md_config_default(md_config *cfg);
const char *md_last_error(void);
const char *md_version(void);
int md_device_count(int *count);

/* ---- context: one GPU, one HIP stream ------------------------------------------------ */
typedef struct md_ctx md_ctx;
int md_ctx_create(int device, md_ctx **out);
int md_ctx_destroy(md_ctx *ctx);
int md_ctx_synchronize(md_ctx *ctx);

/* ---- one mesh: the AsRigidAsPossibleDeformation object ------------------------------ */
typedef struct md_mesh md_mesh;

/* Replaces the constructor + initializeMeshTopology (inc/deform/arap.h:66-70,149-155):
 * n vertices, F faces given as 3F vertex indices. Positions are read at deform time. */
int md_mesh_create(md_ctx *ctx, int n, int F, const int *tri, const md_config *cfg, md_mesh **out);
int md_mesh_destroy(md_mesh *mesh);

/* Replaces setConstraint (inc/deform/arap.h:81-85): insert or overwrite a target and mark
 * the object dirty. Indices outside [0, n) are MD_ERR_INVALID (undefined in the reference). */
int md_mesh_set_constraint(md_mesh *mesh, int vidx, const double *xyz3);
int md_mesh_set_constraints(md_mesh *mesh, int k, const int *idx, const double *xyz);

/* Replaces deform (inc/deform/arap.h:101-138). xyz is the adapter's vertex storage (3n
 * doubles, original vertex order): read when the object is dirty, overwritten by the
 * write-back. *ok receives the reference's bool: 1, or 0 when the factorization failed
 * (then the status is MD_ERR_FACTOR and xyz is untouched). With no constraints it returns
 * with *ok = 1 and leaves xyz untouched, like arap.h:113-114. */
int md_mesh_deform(md_mesh *mesh, int iterations, double *xyz, int *ok);

/* Device-resident variants for measurement and batching: prepare runs the dirty setup of
 * deform() (positions from xyz) without iterating; iterate runs iterations on the device
 * state without any host copy; download copies the current positions (original order). */
int md_mesh_prepare(md_mesh *mesh, const double *xyz);
/* md_mesh_prepare for count meshes at once (positions xyz[k]): the host analysis and
 * factorization of different meshes run on up to nthreads threads (0 = all cores, at most
 * 16); device work stays on the context's stream. */
int md_meshes_prepare(md_ctx *ctx, int count, md_mesh *const *meshes, const double *const *xyz, int nthreads);
int md_mesh_iterate(md_mesh *mesh, int iterations);
int md_mesh_download(md_mesh *mesh, double *xyz);

/* The reference's PrivateAccessor hook (tests/accessor.h:16-22): the cotan weight matrix
 * of the current rest pose as CSR (original numbering, ascending columns, no diagonal). */
int md_mesh_cotan_nnz(md_mesh *mesh, int64_t *nnz);
int md_mesh_cotan_csr(md_mesh *mesh, int *rowptr, int *col, double *val);

/* Rotations of the last local step, row-major 3x3 per vertex in original order. */
int md_mesh_rotations(md_mesh *mesh, double *R);
/* Unit vertex normals of the current positions (3n, original order): the normalised sum of
 * the incident faces' area-weighted normals, for rendering clients that update normals after
 * every deform (examples/osg_viewer.cpp:48). */
int md_mesh_normals(md_mesh *mesh, double *nrm);
/* md_mesh_normals without any allocation: no counterpart in the reference. For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). The face normals and the staging of the result live in one static
 * device workspace per GPU of 64 MiB (held from the library's load until the process ends),
 * enough for 24 F + 60 n bytes (256-byte aligned parts), e.g. 500k vertices with 1M faces;
 * larger meshes return MD_ERR_INVALID (md_mesh_normals has no limit). The mesh must be prepared
 * and resident (an evicted mesh returns MD_ERR_STATE: restoring it would allocate). Every call
 * on one GPU shares that workspace: calls that overlap on the same GPU (from several threads or
 * contexts) race and must be serialised by the caller. Same result as md_mesh_normals. */
int md_mesh_normals_noalloc(md_mesh *mesh, double *nrm);
/* ARAP energy sum_i sum_j w_ij |(p'_i - p'_j) - R_i (p_i - p_j)|^2 of the current state. */
int md_mesh_energy(md_mesh *mesh, double *energy);
int md_mesh_get_stats(md_mesh *mesh, md_mesh_stats *out);

/* Times `iterations` device iterations (after `warmup` untimed ones) with HIP events on the
 * context's stream: total milliseconds, and per-phase milliseconds summed over the timed
 * iterations: [0] local step, [1] residual, [2] forward sweeps, [3] backward sweeps. */
int md_mesh_time_iterations(md_mesh *mesh, int warmup, int iterations, double *total_ms,
                            double *phase_ms4);

/* Autotuner of the sweep layout (MD_SOLVER_LDLT, prepared mesh): for each of the `count`
 * candidate bottom-subtree budgets (md_config.subtree_budget, each in [0, 1024]) the factor
 * is re-laid out on the device and `iterations` iterations are timed with HIP events; the
 * fastest candidate is kept (also in the mesh's config, for later re-preparations). The
 * positions and rotations are restored after every trial, so the mesh state is unchanged.
 * best_budget and best_ms (per iteration) may be NULL. Batches built from the mesh before
 * the call must be rebuilt; a mesh registered with a server is MD_ERR_STATE. */
int md_mesh_autotune(md_mesh *mesh, const int *budgets, int count, int iterations, int *best_budget,
                     double *best_ms);
/* Autotuner of the factor layout as well: for each of the nratios amalgamation ratios
 * (md_config.amalg_ratio) the dirty setup is re-run from the rest pose (analysis and FP64
 * factorization) and md_mesh_autotune chooses the subtree budget; the fastest pair is kept and
 * the mesh is left freshly prepared with it. Call it right after md_mesh_prepare, before any
 * iteration (MD_ERR_STATE otherwise). Costs one setup per ratio. */
int md_mesh_autotune_factor(md_mesh *mesh, const double *ratios, int nratios, const int *budgets,
                            int nbudgets, int iterations, double *best_ratio, int *best_budget,
                            double *best_ms);

/* ---- batches: many prepared meshes iterated in the same launches -------------------- */
typedef struct md_batch md_batch;
/* A batch launches the device arrays its meshes have at creation: a mesh re-prepared with a
 * new constraint set, or re-laid out by md_mesh_autotune, makes the batch's calls return
 * MD_ERR_STATE until it is rebuilt. */
int md_batch_create(md_ctx *ctx, int count, md_mesh *const *meshes, md_batch **out);
int md_batch_iterate(md_batch *batch, int iterations);
int md_batch_time_iterations(md_batch *batch, int warmup, int iterations, double *total_ms);
/* md_mesh_autotune for the whole batch: one budget for all its meshes, chosen by timing
 * batched iterations. */
int md_batch_autotune(md_batch *batch, const int *budgets, int count, int iterations, int *best_budget,
                      double *best_ms);
int md_batch_destroy(md_batch *batch);

/* ---- serving: resident meshes, continuous batching of drag requests ----------------- */
/* A server iterates the admitted drag requests of many meshes in the same launches (one ARAP
 * iteration per boundary for the whole set) and admits queued requests at every iteration
 * boundary (one request in flight per mesh, at most max_active in flight). HBM budget: a
 * mesh's device footprint (md_mesh_stats.device_bytes) counts against hbm_budget_bytes, and a
 * registration that would exceed it is refused (every served mesh stays resident; eviction
 * of idle meshes to host memory is switched off). Requests on one mesh run in submission
 * order, each warm-started from the previous result. */
typedef struct md_server md_server;
typedef struct {
    size_t hbm_budget_bytes; /* device bytes the resident meshes may hold (0 = unlimited) */
    int max_active;          /* requests iterated together (0 = unlimited) */
} md_server_config;

typedef struct {
    uint64_t submitted, completed, failed;
    uint64_t iterations;      /* launches: one ARAP iteration of every request in the set */
    uint64_t mesh_iterations; /* ARAP iterations summed over the requests of every launch */
    uint64_t evictions, restores;
    size_t resident_bytes, peak_resident_bytes;
    int meshes, resident_meshes, queued;
} md_server_stats;

int md_server_create(md_ctx *ctx, const md_server_config *cfg, md_server **out);
/* Registers a prepared mesh (md_mesh_prepare done, MD_REST_FIXED). MD_ERR_BUDGET when the
 * registered meshes' footprint would exceed the budget. The server takes no ownership:
 * destroy the server before the mesh. */
int md_server_add_mesh(md_server *srv, md_mesh *mesh, int *mesh_id);
/* Queues a drag: move the k constrained vertices idx to xyz, then run `iterations`
 * warm-started iterations. Indices must already be constrained on that mesh; a vertex listed
 * more than once takes its last target. */
int md_server_submit(md_server *srv, int mesh_id, int k, const int *idx, const double *xyz,
                     int iterations, uint64_t *ticket);
/* done = 1 when finished; positions (3n, original order, may be NULL) and the latency in
 * milliseconds from submit to completion are returned once done. */
int md_server_poll(md_server *srv, uint64_t ticket, int *done, double *xyz, double *latency_ms);
/* poll and wait hand a finished request's result out once: a second call for the ticket (or
 * a wait whose ticket another call took meanwhile) is MD_ERR_INVALID. */
int md_server_wait(md_server *srv, uint64_t ticket, double timeout_ms, double *xyz, double *latency_ms);
int md_server_get_stats(md_server *srv, md_server_stats *out);
/* Waits for the launches in flight; requests still queued fail with MD_ERR_STATE. */
int md_server_destroy(md_server *srv);

/* ---- device diagnostics ------------------------------------------------------------- */
/* Copies bytes (rounded down to 16) between two fresh device buffers `repeats` times with a
 * plain copy kernel and returns the achieved rate (read + write bytes per second, GB/s):
 * the known-size reference for HBM traffic counters and the achievable copy bandwidth.
 * No counterpart in the reference. */
int md_copy_probe(md_ctx *ctx, size_t bytes, int repeats, double *gbps);

/* ---- host-only diagnostics (no GPU needed) ------------------------------------------ */
/* Factors the SPD matrix given as full CSR (n rows, ascending columns), with coordinates
 * xyz (3 per row) for the ordering, by the same host analysis and FP64 block-LDL^T the
 * library uploads, then solves nrhs right-hand sides (b[r*n + i]) on the host by an
 * emulation of the device sweeps (subtree_budget and level_pairs as in md_config; negative
 * subtree_budget: the default).
 * device_factor_emulation != 0 factors by a host emulation of the device factorization (the
 * same steps and index math as its kernels, every index checked) instead. */
int md_host_factor_solve(int n, const int *rowptr, const int *col, const double *val,
                         const double *xyz, int nrhs, const double *b, double *x,
                         int64_t *panel_entries, int *levels, int subtree_budget,
                         int device_factor_emulation, int level_pairs);

/* The local step's rotation (arap.h:376-382) for count covariances cov (row-major 3x3, 9
 * per entry: sum_j w_ij (p_i - p_j)(p'_i - p'_j)^T) by the kernel's own code compiled for the
 * host: unit quaternions (x, y, z, w) with w >= 0 into quat (4 per entry). prev (4 per entry,
 * may be NULL: identity) is the previous rotation, the warm start of the Jacobi fallback and
 * the result for a zero covariance. fp64 = 0 runs it in FP32, as MD_PREC_FP32_REFINE does.
 * fallback (may be NULL) receives 1 per entry where the top eigenvalue of Horn's matrix was
 * too close to the next for the characteristic-polynomial path and Jacobi sweeps decided. */
int md_host_best_rotation(int fp64, int count, const double *cov, const double *prev, double *quat, int *fallback);

/* One admission pass of the server's planner (the code md_server_* runs at every iteration
 * boundary) on a described state, without a device: nmeshes served meshes with their device
 * footprint, least-recently-used stamp, residency and busy flags; nqueue queued requests given
 * by their mesh; `active` requests in flight, max_active (<= 0: no limit) and the HBM budget
 * (0: no limit); *resident_bytes in and out. Out: take[q] = 1 for admitted requests, the meshes
 * to evict (evict[0 .. *nevict)) and then to restore (restore[0 .. *nrestore)); evict and
 * restore hold nmeshes entries. */
int md_host_admission_plan(int nmeshes, const int64_t *bytes, const uint64_t *last_use, const int *resident,
                           const int *busy, int nqueue, const int *queue, int active, int max_active,
                           int64_t budget, int64_t *resident_bytes, int *take, int *evict, int *nevict,
                           int *restore, int *nrestore);

#ifdef __cplusplus
}
#endif

#endif /* MD_ARAP_H */
