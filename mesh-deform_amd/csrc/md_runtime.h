// Host runtime objects behind the C ABI (include/md_arap.h).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <string>
#include <vector>

#include "md_analysis.h"
#include "md_arap.h"
#include "md_kernels.h"

namespace md {

void set_error(const std::string &msg);

// One device allocation carved into 256-byte aligned buffers.
class Arena {
public:
    ~Arena() { release(); }
    // two passes: reserve sizes, then allocate once and hand out pointers
    size_t reserve(size_t bytes) {
        size_t off = (size_ + 255) & ~size_t(255);
        size_ = off + bytes;
        return off;
    }
    hipError_t allocate() {
        release();
        if (size_ == 0) return hipSuccess;
        return hipMalloc(&base_, size_);
    }
    void *at(size_t off) const { return static_cast<char *>(base_) + off; }
    char *base() const { return static_cast<char *>(base_); }
    bool allocated() const { return base_ != nullptr; }
    size_t size() const { return size_; }
    void release() {
        if (base_) hipFree(base_);
        base_ = nullptr;
    }
    void reset() { release(); size_ = 0; }

private:
    void *base_ = nullptr;
    size_t size_ = 0;
};

// Prefix tables and device mesh table for one set of meshes launched together.
struct Plan {
    int nactive = 0;
    int nlevels = 0;
    int blk_face = 0, blk_slice = 0, blk_resid = 0, blk_ident = 0, blk_apply = 0, blk_sub = 0;
    int blk_subb = 0, off_subb = 0;    // k_sub_bwd: subtrees, then the p' update of level rows
    std::vector<int> blk_fwd, blk_bwd;
    int off_face = 0, off_slice = 0, off_resid = 0, off_ident = 0, off_apply = 0, off_sub = 0;
    int sub_lds_fwd = 0, sub_lds_bwd = 0;  // LDS vector entries of the largest bottom subtree
    int sub_nodes = 0;                 // supernode records of the largest bottom subtree
    std::vector<int> off_fwd, off_bwd;
    MeshDev *d_meshes = nullptr;       // views into d_tables: the mesh records,
    int *d_prefix = nullptr;           //   the prefix tables,
    int *d_tbase = nullptr;            //   per table and mesh: first tile of the level (level sweeps)
    char *d_tables = nullptr;          // one device buffer for the three, uploaded in one copy
    char *h_tables = nullptr;          //   from this pinned mirror
    size_t cap_tables = 0;
    int *d_bmesh = nullptr;            // block -> mesh maps of every table (static batches)
    size_t cap_bmesh = 0;
    std::vector<int> h_bmesh;
    std::vector<size_t> bmesh_off;     // per table: offset into d_bmesh (SIZE_MAX: no map)
    std::vector<int> h_prefix, h_tbase;
    std::vector<MeshDev> h_meshes;
    bool fp64 = true;
    int refine_steps = 1;
    bool pcg = false;
    int pcg_iters = 0;
    hipGraphExec_t graph = nullptr;        // one iteration
    hipGraphExec_t graph_multi = nullptr;  // kGraphIterations iterations
    ~Plan();
    void drop_graph();
    LaunchSet set(int off) const {
        size_t t = (size_t)off / (size_t)(nactive + 1);
        const int *bm = (d_bmesh && t < bmesh_off.size() && bmesh_off[t] != (size_t)-1) ? d_bmesh + bmesh_off[t] : nullptr;
        return LaunchSet{d_meshes, d_prefix + off, nactive, bm, d_tbase + off, nullptr};
    }
    // Level launch: a one-mesh plan also passes the level's first tile, so that blocks load
    // their tile without first loading the mesh record.
    LaunchSet level_set(int off, bool fwd) const {
        LaunchSet L = set(off);
        if (nactive == 1 && !h_meshes.empty())
            L.tiles = (fwd ? h_meshes[0].ftile : h_meshes[0].btile) + h_tbase[(size_t)off];
        return L;
    }
};

// Builds (or rebuilds) a plan for the given meshes and uploads its tables on `stream`.
// With sync = false the uploads are only stream-ordered; the plan's host tables must then
// stay unchanged until the stream has passed them.
hipError_t build_plan(Plan &plan, const std::vector<const MeshDev *> &meshes,
                      const std::vector<const Factor *> &factors, bool fp64, int refine_steps,
                      hipStream_t stream, bool sync = true, bool block_maps = false);
// Enqueues one ARAP iteration (local step + refine_steps x (residual, sweeps)).
void enqueue_iteration(const Plan &plan, hipStream_t stream, hipEvent_t *phase_events);
// Runs `iters` iterations, from a captured graph when use_graph is set.
hipError_t run_iterations(Plan &plan, hipStream_t stream, int iters, bool use_graph);

// Sweep layout of a prepared LDLT mesh with another bottom-subtree budget (device_layout +
// a new layout arena + the mesh's plan). Invalidates plans built from the old layout.
int relayout(md_mesh *m, int budget);

// Device copies of the iteration state (p' and rotations) of meshes, for trial runs.
struct StateSnapshot {
    std::vector<void *> buf;
    ~StateSnapshot();
};
hipError_t snapshot_state(const std::vector<md_mesh *> &meshes, hipStream_t st, StateSnapshot &s);
hipError_t restore_state(const std::vector<md_mesh *> &meshes, hipStream_t st, const StateSnapshot &s);

// Event-timed replay of `iters` iterations of a plan (ms for all of them).
hipError_t time_plan(Plan &plan, hipStream_t st, int iters, bool use_graph, double &ms);

// Pinned host copy of a mesh's two device arenas while the mesh is not resident.
struct HostImage {
    void *a = nullptr, *l = nullptr;   // main arena, sweep-layout arena
    size_t asz = 0, lsz = 0;
    bool complete = false;             // every byte is current (then an eviction copies only
                                       //   the iteration state: p' and the rotations)
    void release();
    ~HostImage() { release(); }
};

// Device residency of a prepared mesh. mesh_evict copies the device state to the host image
// (stream-ordered on st, then synchronised) and frees both arenas; mesh_restore allocates them
// again, copies the image back and re-points every device pointer of the mesh, its launch plan
// included (the generation is bumped: batches built before must be rebuilt). Iterating a
// restored mesh gives bitwise what iterating it without the round trip gives.
int mesh_evict(md_mesh *m, hipStream_t st);
int mesh_restore(md_mesh *m, hipStream_t st);
// Restores an evicted mesh on its context's stream (entry points that touch device state).
int ensure_resident(md_mesh *m, const char *who);

}  // namespace md

struct md_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string arch;
    void *scratch = nullptr;  // this device's address of the static workspace (kScratchBytes)
    // the caller's handle plus one reference per mesh, batch and server created from it:
    // the stream is destroyed with the last holder, whatever the order of destruction
    std::atomic<int> refs{1};
};

namespace md {
void ctx_retain(md_ctx *c);
void ctx_release(md_ctx *c);
}  // namespace md

struct md_mesh {
    md_ctx *ctx = nullptr;
    md_config cfg{};
    int n = 0, F = 0;
    std::vector<int> tri;          // original faces
    std::vector<char> is_con;      // constrained flags (the key set of the reference's map)
    std::vector<double> con;       // targets, 3 per vertex
    int ncon = 0;
    bool dirty = true, con_set_changed = true, initialized = false;
    bool analyzed = false, factored = false, iterated = false;
    std::vector<double> rest;      // rest positions, original order
    md::Topology topo;
    md::Factor fac;
    md::Arena arena;               // everything that does not depend on the sweep layout
    md::Arena larena;              // sweep layout: partials, contribution lists, tiles, subtrees
    md::MeshDev dev{};
    uint64_t gen = 0;              // bumped whenever device pointers in dev change
    // device numeric factorization: static structure (main arena)
    const md::FacDev *d_fac = nullptr;
    const int *d_hlist = nullptr, *d_offrows = nullptr, *d_rel = nullptr;
    double factor_ms = 0.0;        // last numeric factorization (host or device), wall ms
    double deform_ms[6] = {0, 0, 0, 0, 0, 0};  // phases of the last md_mesh_deform (md_mesh_stats)
    double setup_phase_ms[3] = {0, 0, 0};      // phases of the last setup: analysis, uploads +
                                               //   cotan weights, factorization
    std::atomic<int> servers{0};   // servers this mesh is registered with: its device state
                                   //   belongs to the server (entry points that change or read
                                   //   it return MD_ERR_STATE until the server is destroyed)
    bool evicted = false;          // device arenas freed, state held in img
    md::HostImage img;
    const char *img_old_a = nullptr, *img_old_l = nullptr;  // arena addresses at the eviction
    md::Plan plan;                 // single-mesh launch plan
    std::vector<int> adj_ptr, adj; // free-free pattern in solve order
    std::vector<int64_t> adj_slot; // slot of each adj entry
    double setup_ms_host = 0.0;
    std::vector<double> w_host;    // cotan weights on the host (factorization input)
    std::vector<char> pc_host, pr_host;  // device panel layouts staged on the host
    double *d_tgt_xyz = nullptr;   // targets of the constrained ids, internal order (3 (n - nf))
    const int *d_orig = nullptr;   // original id of each internal id (topo.original_of)
    double *d_xyz = nullptr;       // packed xyz in original order (3n): positions in and out
    std::vector<double> tgt_stage; // targets of the constrained ids in internal order (host
                                   //   source of an upload; kept until the stream has passed it)
    // FP64 workspace of the device factorization, kept between calls by meshes that refactor
    // on every deform call (reference rest mode); others free it after each factorization
    void *fac_ws = nullptr;
    size_t fac_ws_bytes = 0;
    void release_fac_ws() {
        if (fac_ws) hipFree(fac_ws);
        fac_ws = nullptr;
        fac_ws_bytes = 0;
    }
    // pinned staging of the packed positions for meshes that download them on every call (from
    // the second download on): chunked copies whose DMA overlaps the host copy to the caller's
    // buffer; pin_ev[c] marks the arrival of chunk c
    static constexpr int kPinChunks = 8;
    double *xyz_pin = nullptr;
    size_t xyz_pin_bytes = 0;
    hipEvent_t pin_ev[kPinChunks] = {};
    int downloads = 0;
    void release_pin() {
        if (xyz_pin) hipHostFree(xyz_pin);
        xyz_pin = nullptr;
        xyz_pin_bytes = 0;
        for (auto &ev : pin_ev)
            if (ev) {
                hipEventDestroy(ev);
                ev = nullptr;
            }
    }
    bool fp64() const { return cfg.precision == MD_PREC_FP64; }
    ~md_mesh() {
        release_fac_ws();
        release_pin();
    }
};
