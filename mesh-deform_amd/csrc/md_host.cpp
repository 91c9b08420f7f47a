// Host-only diagnostic of the C ABI (no HIP): the library's ordering, symbolic analysis and
// FP64 block LDL^T, solved by the bounds-checked emulation of the device sweeps. Compiled by
// the host compiler alone, so that a sanitizer build of the analysis can link it.
#include <algorithm>
#include <string>
#include <vector>

#include "md_admission.h"
#include "md_analysis.h"
#include "md_arap.h"

namespace md {
void set_error(const std::string &msg);
}

using namespace md;

extern "C" int md_host_factor_solve(int n, const int *rowptr, const int *col, const double *val,
                                    const double *xyz, int nrhs, const double *b, double *x,
                                    int64_t *panel_entries, int *levels, int subtree_budget,
                                    int device_factor_emulation, int level_pairs) {
    if (n <= 0 || !rowptr || !col || !val || !xyz || nrhs < 0 || (nrhs > 0 && (!b || !x))) {
        set_error("md_host_factor_solve: invalid argument");
        return MD_ERR_INVALID;
    }
    const int budget = subtree_budget < 0 ? kSubBudgetDefault : std::min(subtree_budget, kSubBudgetMax);
    std::vector<int> perm = order_graph(n, rowptr, col, xyz);
    std::vector<int> group;
    if (level_pairs) {
        std::vector<int> ptr0(n + 1, 0), adj0;
        for (int i = 0; i < n; ++i) {
            for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
                if (col[e] != i) adj0.push_back(col[e]);
            ptr0[i + 1] = (int)adj0.size();
        }
        perm = pair_levels(perm, ptr0, adj0, kPairBudget, kAmalgRatioDefault, group);
    }
    std::vector<int> iperm(n);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    std::vector<int> ptr(n + 1, 0), adj;
    std::vector<double> av, diag(n, 0.0);
    for (int k = 0; k < n; ++k) {
        int i = perm[k];
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            int j = col[e];
            if (j == i) diag[k] += val[e];
            else { adj.push_back(iperm[j]); av.push_back(val[e]); }
        }
        ptr[k + 1] = (int)adj.size();
    }
    Factor f;
    symbolic(n, ptr, adj, f, budget, kAmalgRatioDefault, level_pairs ? &group : nullptr);
    std::string err;
    const bool ok = device_factor_emulation ? numeric_device_emulation(ptr, adj, av, diag, f, err)
                                            : numeric(ptr, adj, av, diag, f, err);
    if (!ok) { set_error(err); return err.rfind("factor emulation", 0) == 0 ? MD_ERR_INVALID : MD_ERR_FACTOR; }
    std::vector<double> bp((size_t)n * nrhs), xp((size_t)n * nrhs);
    for (int r = 0; r < nrhs; ++r)
        for (int k = 0; k < n; ++k) bp[(size_t)r * n + k] = b[(size_t)r * n + perm[k]];
    // three right-hand sides at a time through the tile emulation of the device sweeps
    for (int r0 = 0; r0 < nrhs; r0 += 3) {
        std::vector<double> b3(3 * (size_t)n, 0.0), x3(3 * (size_t)n, 0.0);
        int cnt = std::min(3, nrhs - r0);
        std::copy(bp.begin() + (size_t)r0 * n, bp.begin() + (size_t)(r0 + cnt) * n, b3.begin());
        if (!solve_tiles_host(f, b3.data(), x3.data(), err)) { set_error(err); return MD_ERR_INVALID; }
        std::copy(x3.begin(), x3.begin() + (size_t)cnt * n, xp.begin() + (size_t)r0 * n);
    }
    for (int r = 0; r < nrhs; ++r)
        for (int k = 0; k < n; ++k) x[(size_t)r * n + perm[k]] = xp[(size_t)r * n + k];
    if (panel_entries) *panel_entries = f.factor_nnz();
    if (levels) *levels = f.nlevels;
    return MD_OK;
}

namespace md {

void plan_admission(std::vector<AdmitMesh> &meshes, const std::vector<int> &queue, int active, int max_active,
                    size_t budget, size_t &resident_bytes, uint64_t &clock, std::vector<char> &take,
                    std::vector<int> &evict, std::vector<int> &restore) {
    take.assign(queue.size(), 0);
    evict.clear();
    restore.clear();
    bool blocked = false;  // a request waits for memory: later non-resident meshes wait behind it
    for (size_t q = 0; q < queue.size(); ++q) {
        if (max_active > 0 && active >= max_active) break;
        AdmitMesh &sm = meshes[queue[q]];
        if (sm.busy) continue;
        if (!sm.resident) {
            // pending eviction in this pass: restored at a later boundary, like any evicted mesh
            if (std::find(evict.begin(), evict.end(), queue[q]) != evict.end()) continue;
            if (blocked) continue;
            std::vector<int> victims;
            size_t after = resident_bytes;
            if (budget > 0) {
                std::vector<int> idle;
                for (int k = 0; k < (int)meshes.size(); ++k)
                    if (meshes[k].resident && !meshes[k].busy) idle.push_back(k);
                std::stable_sort(idle.begin(), idle.end(),
                                 [&](int a, int b) { return meshes[a].last_use < meshes[b].last_use; });
                for (int v : idle) {
                    if (after + sm.bytes <= budget) break;
                    victims.push_back(v);
                    after -= meshes[v].bytes;
                }
                if (after + sm.bytes > budget) {
                    blocked = true;
                    continue;
                }
            }
            for (int v : victims) {
                meshes[v].resident = false;
                evict.push_back(v);
            }
            sm.resident = true;
            restore.push_back(queue[q]);
            resident_bytes = after + sm.bytes;
        }
        sm.busy = true;
        sm.last_use = ++clock;
        ++active;
        take[q] = 1;
    }
}

}  // namespace md

// Host diagnostic: one admission pass of the server's planner on a described state.
extern "C" int md_host_admission_plan(int nmeshes, const int64_t *bytes, const uint64_t *last_use,
                                      const int *resident, const int *busy, int nqueue, const int *queue,
                                      int active, int max_active, int64_t budget, int64_t *resident_bytes,
                                      int *take, int *evict, int *nevict, int *restore, int *nrestore) {
    if (nmeshes < 0 || nqueue < 0 || budget < 0 || !resident_bytes || !nevict || !nrestore ||
        (nmeshes > 0 && (!bytes || !last_use || !resident || !busy || !evict || !restore)) ||
        (nqueue > 0 && (!queue || !take))) {
        set_error("md_host_admission_plan: invalid argument");
        return MD_ERR_INVALID;
    }
    std::vector<AdmitMesh> ms(nmeshes);
    for (int k = 0; k < nmeshes; ++k) {
        if (bytes[k] < 0) { set_error("md_host_admission_plan: negative footprint"); return MD_ERR_INVALID; }
        ms[k].bytes = (size_t)bytes[k];
        ms[k].last_use = last_use[k];
        ms[k].resident = resident[k] != 0;
        ms[k].busy = busy[k] != 0;
    }
    std::vector<int> qv(queue, queue + nqueue);
    for (int m : qv)
        if (m < 0 || m >= nmeshes) { set_error("md_host_admission_plan: queued request on an unknown mesh"); return MD_ERR_INVALID; }
    size_t rb = (size_t)*resident_bytes;
    uint64_t clock = 0;
    for (int k = 0; k < nmeshes; ++k) clock = std::max(clock, ms[k].last_use);
    std::vector<char> tk;
    std::vector<int> ev, rs;
    plan_admission(ms, qv, active, max_active, (size_t)budget, rb, clock, tk, ev, rs);
    for (int q = 0; q < nqueue; ++q) take[q] = tk[q];
    for (size_t i = 0; i < ev.size(); ++i) evict[i] = ev[i];
    for (size_t i = 0; i < rs.size(); ++i) restore[i] = rs[i];
    *nevict = (int)ev.size();
    *nrestore = (int)rs.size();
    *resident_bytes = (int64_t)rb;
    return MD_OK;
}
