// Host-only diagnostic of the C ABI (no HIP): the library's ordering, symbolic analysis and
// FP64 block LDL^T, solved by the bounds-checked emulation of the device sweeps. Compiled by
// the host compiler alone, so that a sanitizer build of the analysis can link it.
#include <algorithm>
#include <string>
#include <vector>

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
