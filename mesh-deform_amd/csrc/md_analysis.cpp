// Host-side analysis for the ARAP solve: see md_analysis.h.
#include "md_analysis.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace md {

namespace {

// Free-free adjacency in "free rank" numbering (rank among free vertices, ascending id).
void free_graph(int n, int F, const int* tri, const std::vector<char>& is_con,
                std::vector<int>& rank_of, std::vector<int>& vert_of, std::vector<int>& ptr,
                std::vector<int>& adj) {
    rank_of.assign(n, -1);
    vert_of.clear();
    for (int v = 0; v < n; ++v)
        if (!is_con[v]) { rank_of[v] = (int)vert_of.size(); vert_of.push_back(v); }
    int m = (int)vert_of.size();
    std::vector<std::pair<int, int>> e;
    e.reserve((size_t)F * 6);
    for (int f = 0; f < F; ++f)
        for (int k = 0; k < 3; ++k) {
            int a = rank_of[tri[3 * f + k]], b = rank_of[tri[3 * f + (k + 1) % 3]];
            if (a >= 0 && b >= 0 && a != b) { e.emplace_back(a, b); e.emplace_back(b, a); }
        }
    std::sort(e.begin(), e.end());
    e.erase(std::unique(e.begin(), e.end()), e.end());
    ptr.assign(m + 1, 0);
    adj.resize(e.size());
    for (size_t k = 0; k < e.size(); ++k) { ptr[e[k].first + 1]++; adj[k] = e[k].second; }
    for (int i = 0; i < m; ++i) ptr[i + 1] += ptr[i];
}

// ---- geometric nested dissection ------------------------------------------------------

struct ND {
    const std::vector<int>& ptr;
    const std::vector<int>& adj;
    const std::vector<double>& xyz;  // 3 per free rank
    std::vector<int> stamp, side;
    int stamp_ctr = 0;
    std::vector<int> out;
    static constexpr int kLeaf = 8;

    ND(const std::vector<int>& p, const std::vector<int>& a, const std::vector<double>& x)
        : ptr(p), adj(a), xyz(x), stamp(p.size(), 0), side(p.size(), 0) {}

    void run(std::vector<int>& v) {
        int cnt = (int)v.size();
        if (cnt <= kLeaf) {
            out.insert(out.end(), v.begin(), v.end());
            return;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int x : v)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], xyz[3 * x + a]);
                hi[a] = std::max(hi[a], xyz[3 * x + a]);
            }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
        int half = cnt / 2;
        std::nth_element(v.begin(), v.begin() + half, v.end(), [&](int i, int j) {
            double a = xyz[3 * i + axis], b = xyz[3 * j + axis];
            return a < b || (a == b && i < j);
        });
        int st = ++stamp_ctr;
        for (int k = 0; k < cnt; ++k) { stamp[v[k]] = st; side[v[k]] = k < half ? 0 : 1; }
        // boundary of each side; the smaller one is the separator
        std::vector<int> bnd[2], rest[2];
        for (int k = 0; k < cnt; ++k) {
            int x = v[k], sx = side[x];
            bool b = false;
            for (int e = ptr[x]; e < ptr[x + 1]; ++e) {
                int y = adj[e];
                if (stamp[y] == st && side[y] != sx) { b = true; break; }
            }
            (b ? bnd[sx] : rest[sx]).push_back(x);
        }
        int ss = bnd[0].size() <= bnd[1].size() ? 0 : 1;
        std::vector<int> left, right, sep = bnd[ss];
        left = rest[0];
        right = rest[1];
        if (ss == 0) right.insert(right.end(), bnd[1].begin(), bnd[1].end());
        else left.insert(left.end(), bnd[0].begin(), bnd[0].end());
        std::vector<int>().swap(v);
        run(left);
        run(right);
        out.insert(out.end(), sep.begin(), sep.end());
    }
};

// Elimination tree of the symmetric pattern (ptr/adj) under the order `perm` (new -> old).
std::vector<int> etree(const std::vector<int>& ptr, const std::vector<int>& adj,
                       const std::vector<int>& perm, const std::vector<int>& iperm) {
    int m = (int)perm.size();
    std::vector<int> parent(m, -1), anc(m, -1);
    for (int k = 0; k < m; ++k) {
        int old = perm[k];
        for (int e = ptr[old]; e < ptr[old + 1]; ++e) {
            int i = iperm[adj[e]];
            while (i != -1 && i < k) {
                int nxt = anc[i];
                anc[i] = k;
                if (nxt == -1) { parent[i] = k; break; }
                i = nxt;
            }
        }
    }
    return parent;
}

std::vector<int> postorder(const std::vector<int>& parent) {
    int m = (int)parent.size();
    std::vector<int> head(m, -1), next(m, -1), post;
    post.reserve(m);
    for (int j = m - 1; j >= 0; --j)
        if (parent[j] != -1) { next[j] = head[parent[j]]; head[parent[j]] = j; }
    std::vector<int> stack;
    for (int r = 0; r < m; ++r) {
        if (parent[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            int p = stack.back();
            int c = head[p];
            if (c == -1) { stack.pop_back(); post.push_back(p); }
            else { head[p] = next[c]; stack.push_back(c); }
        }
    }
    return post;
}

}  // namespace

namespace {

// Nested dissection of the graph (ptr/adj, m vertices with coordinates fx) followed by an
// elimination-tree postorder; returns new -> old.
std::vector<int> nd_postorder(int m, const std::vector<int>& ptr, const std::vector<int>& adj,
                              const std::vector<double>& fx) {
    ND nd(ptr, adj, fx);
    std::vector<int> all(m);
    std::iota(all.begin(), all.end(), 0);
    nd.out.reserve(m);
    nd.run(all);
    std::vector<int> perm = nd.out, iperm(m);
    for (int k = 0; k < m; ++k) iperm[perm[k]] = k;
    // postorder the elimination tree: same fill, contiguous subtrees
    std::vector<int> par = etree(ptr, adj, perm, iperm);
    std::vector<int> post = postorder(par);
    std::vector<int> result(m);
    for (int k = 0; k < m; ++k) result[k] = perm[post[k]];
    return result;
}

}  // namespace

std::vector<int> order_graph(int m, const int* rowptr, const int* col, const double* xyz) {
    std::vector<int> ptr(m + 1, 0), adj;
    for (int i = 0; i < m; ++i) {
        for (int e = rowptr[i]; e < rowptr[i + 1]; ++e)
            if (col[e] != i) adj.push_back(col[e]);
        ptr[i + 1] = (int)adj.size();
    }
    std::vector<double> fx(xyz, xyz + 3 * (size_t)m);
    return nd_postorder(m, ptr, adj, fx);
}

std::vector<int> order_free_vertices(int n, int F, const int* tri, const double* xyz,
                                     const std::vector<char>& is_con, std::vector<int>* pair_group,
                                     double amalg_ratio) {
    std::vector<int> rank_of, vert_of, ptr, adj;
    free_graph(n, F, tri, is_con, rank_of, vert_of, ptr, adj);
    int m = (int)vert_of.size();
    std::vector<double> fx(3 * (size_t)m);
    for (int k = 0; k < m; ++k)
        for (int a = 0; a < 3; ++a) fx[3 * k + a] = xyz[3 * (size_t)vert_of[k] + a];
    std::vector<int> perm = nd_postorder(m, ptr, adj, fx);
    if (pair_group) perm = pair_levels(perm, ptr, adj, kPairBudget, amalg_ratio, *pair_group);
    for (int k = 0; k < m; ++k) perm[k] = vert_of[perm[k]];
    return perm;
}

void build_topology(int n, int F, const int* tri, const std::vector<int>& perm,
                    const std::vector<char>& is_con, Topology& t) {
    t.n = n;
    t.F = F;
    t.nf = (int)perm.size();
    t.internal_of.assign(n, -1);
    t.original_of.assign(n, -1);
    int next = 0;
    for (int v : perm) { t.internal_of[v] = next; t.original_of[next] = v; ++next; }
    for (int v = 0; v < n; ++v)
        if (is_con[v]) { t.internal_of[v] = next; t.original_of[next] = v; ++next; }
    // undirected edge list with the face edges that touch it (3f + k), in face order
    std::vector<int64_t> key;
    key.reserve((size_t)F * 3);
    std::vector<int> fe_order;
    for (int f = 0; f < F; ++f)
        for (int k = 0; k < 3; ++k) {
            int a = tri[3 * f + k], b = tri[3 * f + (k + 1) % 3];
            if (a == b) { key.push_back(-1); continue; }
            int lo = std::min(a, b), hi = std::max(a, b);
            key.push_back((int64_t)lo * n + hi);
        }
    // CSR adjacency in original ids (ascending neighbour ids), with per-directed-edge face lists
    std::vector<int> deg(n, 0);
    std::vector<int64_t> ukeys;
    ukeys.reserve(key.size());
    for (int64_t k : key)
        if (k >= 0) ukeys.push_back(k);
    std::sort(ukeys.begin(), ukeys.end());
    ukeys.erase(std::unique(ukeys.begin(), ukeys.end()), ukeys.end());
    for (int64_t k : ukeys) { deg[k / n]++; deg[k % n]++; }
    std::vector<int> rp(n + 1, 0);
    for (int v = 0; v < n; ++v) rp[v + 1] = rp[v] + deg[v];
    std::vector<int> ci(rp[n]);
    std::vector<int> fill(rp.begin(), rp.end() - 1);
    for (int64_t k : ukeys) {
        int a = (int)(k / n), b = (int)(k % n);
        ci[fill[a]++] = b;
        ci[fill[b]++] = a;
    }
    for (int v = 0; v < n; ++v) std::sort(ci.begin() + rp[v], ci.begin() + rp[v + 1]);
    // slices over internal ids
    t.nslices = (n + kSlice - 1) / kSlice;
    t.slice_ptr.assign(t.nslices + 1, 0);
    for (int s = 0; s < t.nslices; ++s) {
        int wmax = 0;
        for (int l = 0; l < kSlice; ++l) {
            int iv = s * kSlice + l;
            if (iv >= n) break;
            int v = t.original_of[iv];
            wmax = std::max(wmax, rp[v + 1] - rp[v]);
        }
        t.slice_ptr[s + 1] = t.slice_ptr[s] + wmax;
    }
    int64_t slots = t.total_slots();
    t.sell_col.assign(slots, 0);
    // directed edge (original a -> original b) -> slot
    std::vector<int64_t> slot_of_csr(rp[n]);
    for (int s = 0; s < t.nslices; ++s) {
        int wdt = t.slice_ptr[s + 1] - t.slice_ptr[s];
        for (int l = 0; l < kSlice; ++l) {
            int iv = s * kSlice + l;
            for (int k = 0; k < wdt; ++k) {
                int64_t slot = ((int64_t)t.slice_ptr[s] + k) * kSlice + l;
                if (iv >= n) { t.sell_col[slot] = 0; continue; }
                int v = t.original_of[iv];
                int d = rp[v + 1] - rp[v];
                if (k < d) {
                    t.sell_col[slot] = t.internal_of[ci[rp[v] + k]];
                    slot_of_csr[rp[v] + k] = slot;
                } else {
                    t.sell_col[slot] = iv;  // padding: self, zero weight
                }
            }
        }
    }
    // face edges feeding each slot, in ascending face order (deterministic summation that
    // matches setFromTriplets' insertion order, arap.h:224-238)
    std::vector<int> cnt(slots + 1, 0);
    auto csr_find = [&](int a, int b) {
        auto it = std::lower_bound(ci.begin() + rp[a], ci.begin() + rp[a + 1], b);
        return (int)(it - ci.begin());
    };
    std::vector<std::pair<int64_t, int>> pairs;
    pairs.reserve((size_t)F * 6);
    for (int f = 0; f < F; ++f)
        for (int k = 0; k < 3; ++k) {
            int a = tri[3 * f + k], b = tri[3 * f + (k + 1) % 3];
            if (a == b) continue;
            pairs.emplace_back(slot_of_csr[csr_find(a, b)], 3 * f + k);
            pairs.emplace_back(slot_of_csr[csr_find(b, a)], 3 * f + k);
        }
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const std::pair<int64_t, int>& x, const std::pair<int64_t, int>& y) {
                         return x.first < y.first;
                     });
    t.slot_fe_ptr.assign(slots + 1, 0);
    t.slot_fe.resize(pairs.size());
    for (size_t k = 0; k < pairs.size(); ++k) {
        t.slot_fe_ptr[pairs[k].first + 1]++;
        t.slot_fe[k] = pairs[k].second;
    }
    for (int64_t s = 0; s < slots; ++s) t.slot_fe_ptr[s + 1] += t.slot_fe_ptr[s];
    t.tri_internal.resize(3 * (size_t)F);
    for (size_t k = 0; k < 3 * (size_t)F; ++k) t.tri_internal[k] = t.internal_of[tri[k]];
}

// ---- symbolic ---------------------------------------------------------------------------

// Supernodes of at most this many columns always merge with their parent.
constexpr int kAmalgSmall = 16;

void symbolic(int n, const std::vector<int>& ptr, const std::vector<int>& adj, Factor& f, int sub_budget,
              double amalg_ratio, const std::vector<int>* group) {
    f = Factor();
    f.n = n;
    std::vector<int> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    std::vector<int> par = etree(ptr, adj, ident, ident);
    // column counts (with diagonal) by row-subtree traversal
    std::vector<int> cc(n, 1), mark(n, -1);
    for (int i = 0; i < n; ++i) {
        mark[i] = i;
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
            int k = adj[e];
            if (k >= i) continue;
            while (k != -1 && k < i && mark[k] != i) {
                cc[k]++;
                mark[k] = i;
                k = par[k];
            }
        }
    }
    f.nnz_exact = 0;
    for (int j = 0; j < n; ++j) f.nnz_exact += cc[j] - 1;
    std::vector<int> nchild(n, 0);
    for (int j = 0; j < n; ++j)
        if (par[j] != -1) nchild[par[j]]++;
    // fundamental supernodes
    std::vector<int> first;
    for (int j = 0; j < n; ++j) {
        bool join = j > 0 && par[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1 &&
                    (j - first.back()) < kMaxSnWidth;
        if (!join) first.push_back(j);
    }
    int ns = (int)first.size();
    first.push_back(n);
    std::vector<int> sn_of(n);
    for (int s = 0; s < ns; ++s)
        for (int j = first[s]; j < first[s + 1]; ++j) sn_of[j] = s;
    // supernodal structure (off-diagonal rows), children before parents
    std::vector<std::vector<int>> rows(ns);
    std::vector<int> spar(ns, -1);
    std::vector<std::vector<int>> kids(ns);
    std::vector<int> marker(n, -1);
    for (int s = 0; s < ns; ++s) {
        int lo = first[s], hi = first[s + 1];
        std::vector<int>& r = rows[s];
        for (int j = lo; j < hi; ++j)
            for (int e = ptr[j]; e < ptr[j + 1]; ++e) {
                int i = adj[e];
                if (i >= hi && marker[i] != s) { marker[i] = s; r.push_back(i); }
            }
        for (int c : kids[s])
            for (int i : rows[c])
                if (i >= hi && marker[i] != s) { marker[i] = s; r.push_back(i); }
        std::sort(r.begin(), r.end());
        if (!r.empty()) {
            spar[s] = sn_of[r[0]];
            kids[spar[s]].push_back(s);
        }
    }
    // relaxed amalgamation: merge the last child into its parent when the panel grows little
    std::vector<int> lo(ns), w(ns);
    for (int s = 0; s < ns; ++s) { lo[s] = first[s]; w[s] = first[s + 1] - first[s]; }
    std::vector<char> alive(ns, 1);
    // explicit zeros each supernode's panel carries after merges
    std::vector<int64_t> zeros(ns, 0);
    for (int p = 0; p < ns; ++p) {
        for (;;) {
            if (kids[p].empty()) break;
            int c = kids[p].back();  // children are appended in increasing order
            if (lo[c] + w[c] != lo[p]) break;
            int64_t wc = w[c], wp = w[p], nc = (int64_t)rows[c].size(), np_ = (int64_t)rows[p].size();
            int64_t before = wc * (wc + nc) + wp * (wp + np_);
            int64_t after = (wc + wp) * (wc + wp + np_);
            int64_t z = zeros[c] + zeros[p] + (after - before);
            double frac = (double)z / (double)after;
            int64_t wm = wc + wp;
            // a level pair (pair_levels placed its columns together) always merges
            const bool grouped = group && (*group)[lo[c]] >= 0 && (*group)[lo[c]] == (*group)[lo[p]];
            if (wm > (grouped ? kMaxGroupWidth : kMaxSnWidth)) break;
            // relaxed amalgamation: tiny supernodes always merge, wider ones only while the
            // share of explicit zeros stays small
            bool ok = grouped || wm <= kAmalgSmall || (wm <= 16 && frac < 0.5) ||
                      (wm <= 48 && frac < amalg_ratio) || frac < 0.5 * amalg_ratio;
            if (!ok) break;
            zeros[p] = z;
            // merge c into p
            kids[p].pop_back();
            for (int g : kids[c]) { spar[g] = p; }
            kids[p].insert(kids[p].end(), kids[c].begin(), kids[c].end());
            std::sort(kids[p].begin(), kids[p].end());
            lo[p] = lo[c];
            w[p] += w[c];
            alive[c] = 0;
            std::vector<int>().swap(rows[c]);
        }
    }
    // renumber surviving supernodes (column order = topological order)
    std::vector<int> newid(ns, -1);
    int nn = 0;
    for (int s = 0; s < ns; ++s)
        if (alive[s]) newid[s] = nn++;
    f.ns = nn;
    f.c0.resize(nn); f.w.resize(nn); f.noff.resize(nn); f.parent.assign(nn, -1);
    f.roff.assign(nn + 1, 0);
    for (int s = 0; s < ns; ++s) {
        if (!alive[s]) continue;
        int t = newid[s];
        f.c0[t] = lo[s];
        f.w[t] = w[s];
        f.noff[t] = (int)rows[s].size();
        f.parent[t] = spar[s] >= 0 ? newid[spar[s]] : -1;
        f.roff[t + 1] = f.noff[t];
    }
    for (int t = 0; t < nn; ++t) f.roff[t + 1] += f.roff[t];
    f.offrows.resize(f.roff[nn]);
    for (int s = 0; s < ns; ++s)
        if (alive[s]) std::copy(rows[s].begin(), rows[s].end(), f.offrows.begin() + f.roff[newid[s]]);
    // children, heights
    f.child_ptr.assign(nn + 1, 0);
    for (int t = 0; t < nn; ++t)
        if (f.parent[t] >= 0) f.child_ptr[f.parent[t] + 1]++;
    for (int t = 0; t < nn; ++t) f.child_ptr[t + 1] += f.child_ptr[t];
    f.child.resize(f.child_ptr[nn]);
    {
        std::vector<int> fill(f.child_ptr.begin(), f.child_ptr.end() - 1);
        for (int t = 0; t < nn; ++t)
            if (f.parent[t] >= 0) f.child[fill[f.parent[t]]++] = t;
    }
    f.height.assign(nn, 0);
    f.nlevels = nn > 0 ? 1 : 0;
    for (int t = 0; t < nn; ++t) {
        for (int k = f.child_ptr[t]; k < f.child_ptr[t + 1]; ++k)
            f.height[t] = std::max(f.height[t], f.height[f.child[k]] + 1);
        f.nlevels = std::max(f.nlevels, f.height[t] + 1);
    }
    // rel: position of each off row of t in its parent's row list [0, m_p)
    f.rel.assign(f.offrows.size(), -1);
    for (int t = 0; t < nn; ++t) {
        int p = f.parent[t];
        if (p < 0) continue;
        const int* prow = f.offrows.data() + f.roff[p];
        int pn = f.noff[p];
        for (int k = f.roff[t]; k < f.roff[t + 1]; ++k) {
            int r = f.offrows[k];
            if (r < f.c0[p] + f.w[p]) {
                f.rel[k] = r - f.c0[p];
            } else {
                int pos = (int)(std::lower_bound(prow, prow + pn, r) - prow);
                f.rel[k] = f.w[p] + pos;
            }
        }
    }
    f.poff.assign(nn + 1, 0);
    for (int t = 0; t < nn; ++t) f.poff[t + 1] = f.poff[t] + (int64_t)(f.w[t] + f.noff[t]) * f.w[t];
    // device numeric factorization: heights, workspace offsets, column owners
    f.hptr.assign(f.nlevels + 1, 0);
    for (int t = 0; t < nn; ++t) f.hptr[f.height[t] + 1]++;
    for (int h = 0; h < f.nlevels; ++h) f.hptr[h + 1] += f.hptr[h];
    f.hlist.assign(nn, 0);
    {
        std::vector<int> fill(f.hptr.begin(), f.hptr.end() - 1);
        for (int t = 0; t < nn; ++t) f.hlist[fill[f.height[t]]++] = t;
    }
    f.woff_s.assign(nn + 1, 0); f.woff_t.assign(nn + 1, 0); f.woff_u.assign(nn + 1, 0);
    for (int t = 0; t < nn; ++t) {
        const int64_t w = f.w[t], no = f.noff[t];
        f.woff_s[t + 1] = f.woff_s[t] + no * w;
        f.woff_t[t + 1] = f.woff_t[t] + w * w;
        f.woff_u[t + 1] = f.woff_u[t] + no * no;
    }
    f.col_sn.assign(n, 0);
    for (int t = 0; t < nn; ++t)
        for (int k = 0; k < f.w[t]; ++k) f.col_sn[f.c0[t] + k] = t;
    device_layout(f, sub_budget);
}

void device_layout(Factor& f, int sub_budget) {
    const int n = f.n, nn = f.ns;
    auto cdiv = [](int a, int b) { return (a + b - 1) / b; };
    // bottom subtrees: maximal subtrees of small supernodes within the LDS budget
    std::vector<int> sub_root(nn, -1), size(nn, 1);
    std::vector<int64_t> lds(nn, 0);
    std::vector<char> ok(nn, 0);
    // subtree records hold int offsets, and a partial count (< 2^15) with a stride (< 2^16)
    int64_t panel_total = 0, max_kr = 0;
    for (int t = 0; t < nn; ++t) {
        panel_total += (int64_t)(f.w[t] + f.noff[t]) * f.w[t];
        max_kr = std::max<int64_t>(max_kr, cdiv(f.w[t] + f.noff[t], kTile));
    }
    if (sub_budget > 0 && panel_total < (int64_t)INT32_MAX / 2 && max_kr < (1 << 15)) {
        for (int t = 0; t < nn; ++t) {
            int m = f.w[t] + f.noff[t];
            bool good = f.w[t] <= kSubMaxWidth && m <= kSubMaxRows;
            lds[t] = m + f.w[t];
            for (int k = f.child_ptr[t]; k < f.child_ptr[t + 1]; ++k) {
                int c = f.child[k];
                good = good && ok[c];
                lds[t] += lds[c];
                size[t] += size[c];
            }
            ok[t] = good && lds[t] <= sub_budget && size[t] <= kSubMaxNodes;
        }
        for (int t = nn - 1; t >= 0; --t) {
            int p = f.parent[t];
            if (!ok[t]) continue;
            sub_root[t] = (p >= 0 && ok[p]) ? sub_root[p] : t;
        }
    }
    f.level.assign(nn, -1);
    f.nlevels = 0;
    for (int t = 0; t < nn; ++t) {
        if (sub_root[t] >= 0) continue;
        int l = 0;
        for (int k = f.child_ptr[t]; k < f.child_ptr[t + 1]; ++k) {
            int c = f.child[k];
            if (sub_root[c] < 0) l = std::max(l, f.level[c] + 1);
        }
        f.level[t] = l;
        f.nlevels = std::max(f.nlevels, l + 1);
    }
    // device layout: tiles, partial vectors, bucket pointers, row locators
    f.pc_off.assign(nn + 1, 0); f.pr_off.assign(nn + 1, 0);
    f.xoff.assign(nn + 1, 0);
    f.kc.resize(nn); f.kr.resize(nn);
    for (int t = 0; t < nn; ++t) {
        int m = f.w[t] + f.noff[t];
        bool in_sub = sub_root[t] >= 0;
        f.kc[t] = in_sub ? (sub_root[t] == t ? 1 : 0) : cdiv(f.w[t], kTile);
        f.kr[t] = in_sub ? 1 : cdiv(m, kTile);
        f.pc_off[t + 1] = f.pc_off[t] + (int64_t)f.noff[t] * f.w[t];
        f.pr_off[t + 1] = f.pr_off[t] + (int64_t)m * f.w[t];
        f.xoff[t + 1] = f.xoff[t] + (int64_t)f.kr[t] * f.w[t];
    }
    // update slots: per level front position as many slots as its largest count of child
    // partials; each child entry's partials go to consecutive slots in child order
    f.ubase.assign(nn + 1, 0);
    f.uk.assign(nn, 0);
    f.udst.assign(f.offrows.size(), -1);
    {
        std::vector<int> cnt;
        for (int t = 0; t < nn; ++t) {
            f.ubase[t + 1] = f.ubase[t];
            if (sub_root[t] >= 0) continue;  // swept by its subtree
            const int m = f.w[t] + f.noff[t];
            cnt.assign(m, 0);
            for (int k = f.child_ptr[t]; k < f.child_ptr[t + 1]; ++k) {
                const int c = f.child[k];
                for (int e = 0; e < f.noff[c]; ++e) {
                    const int p = f.rel[f.roff[c] + e];
                    f.udst[f.roff[c] + e] = f.ubase[t] + (int64_t)cnt[p] * m + p;
                    cnt[p] += f.kc[c];
                }
            }
            int K = 0;
            for (int p = 0; p < m; ++p) K = std::max(K, cnt[p]);
            f.uk[t] = K;
            f.ubase[t + 1] = f.ubase[t] + (int64_t)K * m;
        }
    }
    f.rowinfo.assign(4 * (size_t)n, 0);
    for (int t = 0; t < nn; ++t)
        for (int q = 0; q < f.w[t]; ++q) {
            int row = f.c0[t] + q;
            f.rowinfo[4 * (size_t)row] = (int)(f.xoff[t] + q);
            f.rowinfo[4 * (size_t)row + 1] = f.w[t];
            f.rowinfo[4 * (size_t)row + 2] = f.kr[t];
        }
    f.xloc.assign(4 * f.offrows.size(), 0);
    for (size_t e = 0; e < f.offrows.size(); ++e)
        for (int k = 0; k < 4; ++k) f.xloc[4 * e + k] = f.rowinfo[4 * (size_t)f.offrows[e] + k];
    f.ft_ptr.assign(f.nlevels + 1, 0);
    f.bt_ptr.assign(f.nlevels + 1, 0);
    // a level supernode without off rows (a root) has no forward tiles: its backward tiles
    // gather its right-hand side and children's contributions themselves
    for (int t = 0; t < nn; ++t) {
        if (f.level[t] < 0) continue;
        int m = f.w[t] + f.noff[t];
        if (f.noff[t] > 0) f.ft_ptr[f.level[t] + 1] += cdiv(f.noff[t], kTile) * f.kc[t];
        f.bt_ptr[f.level[t] + 1] += f.kc[t] * cdiv(m, kTile);
    }
    for (int l = 0; l < f.nlevels; ++l) {
        f.ft_ptr[l + 1] += f.ft_ptr[l];
        f.bt_ptr[l + 1] += f.bt_ptr[l];
    }
    f.ft.assign(4 * (size_t)f.ft_ptr[f.nlevels], 0);
    f.bt.assign(4 * (size_t)f.bt_ptr[f.nlevels], 0);
    std::vector<int> ff(f.ft_ptr.begin(), f.ft_ptr.end() - 1), bf(f.bt_ptr.begin(), f.bt_ptr.end() - 1);
    for (int t = 0; t < nn; ++t) {
        int l = f.level[t], m = f.w[t] + f.noff[t];
        if (l < 0) continue;
        int rbs = cdiv(f.noff[t], kTile);
        for (int rb = 0; rb < rbs; ++rb)
            for (int cb = 0; cb < f.kc[t]; ++cb) {
                int* tt = &f.ft[4 * (size_t)ff[l]++];
                tt[0] = t; tt[1] = rb * kTile; tt[2] = cb * kTile;
            }
        for (int cb = 0; cb < f.kc[t]; ++cb)
            for (int pb = 0; pb < cdiv(m, kTile); ++pb) {
                int* tt = &f.bt[4 * (size_t)bf[l]++];
                tt[0] = t; tt[1] = cb * kTile; tt[2] = pb * kTile;
            }
    }
    // bottom subtrees (the subtree of root r is the id range [r - size + 1, r]): per-supernode
    // records by height, contribution entries, backward off-row sources
    f.nsub = 0; f.sub_lds_fwd = 0; f.sub_lds_bwd = 0; f.sub_nodes = 0;
    f.sub.clear(); f.sub_h.clear(); f.snode.clear(); f.scp.clear(); f.sfc.clear(); f.sbo.clear();
    f.srow.clear();
    f.sfi.assign(n, -1);
    std::vector<int> fb(nn, 0), fbb(nn, 0);
    for (int r = 0; r < nn; ++r) {
        if (sub_root[r] != r) continue;
        const int first = r - size[r] + 1, nh = f.height[r] + 1;
        const int col0 = f.c0[first], col1 = f.c0[r] + f.w[r];
        std::vector<std::vector<int>> at(nh);
        for (int t = first; t <= r; ++t) at[f.height[t]].push_back(t);
        // forward: every front; backward: the columns, then the fronts of one height
        int fronts = 0, height_fronts = 0;
        for (int t = first; t <= r; ++t) { fb[t] = fronts; fronts += f.w[t] + f.noff[t]; }
        const int ncols = col1 - col0;
        f.srow.push_back(col0);
        f.srow.push_back(ncols);
        for (int t = first; t <= r; ++t)
            for (int q = 0; q < f.w[t]; ++q) f.sfi[f.c0[t] + q] = fb[t] + q;
        for (int h = 0; h < nh; ++h) {
            int off = ncols;
            for (int t : at[h]) { fbb[t] = off; off += f.w[t] + f.noff[t]; }
            height_fronts = std::max(height_fronts, off - ncols);
        }
        const int rec0 = (int)(f.snode.size() / 16);
        f.sub.push_back(rec0);
        f.sub.push_back(rec0 + size[r]);
        f.sub.push_back((int)f.sub_h.size());
        f.sub.push_back(nh);
        f.sub_lds_fwd = std::max(f.sub_lds_fwd, fronts);
        f.sub_lds_bwd = std::max(f.sub_lds_bwd, ncols + height_fronts);
        f.sub_nodes = std::max(f.sub_nodes, size[r]);
        ++f.nsub;
        std::vector<std::vector<int>> lists;
        for (int h = 0; h <= nh; ++h) {
            f.sub_h.push_back((int)(f.snode.size() / 16));
            if (h == nh) break;
            int pa = 0, pb = 0, pq = 0;
            for (int t : at[h]) {
                const int w = f.w[t], no = f.noff[t], m = w + no;
                int rec[16] = {f.c0[t], w, no, fb[t], (int)f.pc_off[t], (int)f.pr_off[t], (int)(f.scp.size() / 2),
                               (int)(f.sbo.size() / 2), t == r ? f.roff[t] : -1, (int)f.xoff[t],
                               f.c0[t] - col0, pa, pb, pq, fbb[t], 0};
                f.snode.insert(f.snode.end(), rec, rec + 16);
                pa += m; pb += no; pq += w;
                // children's update entries landing on each position, in child order
                lists.assign(m, {});
                for (int k = f.child_ptr[t]; k < f.child_ptr[t + 1]; ++k) {
                    int c = f.child[k];
                    for (int e = 0; e < f.noff[c]; ++e)
                        lists[f.rel[f.roff[c] + e]].push_back(fb[c] + f.w[c] + e);
                }
                for (int p = 0; p < m; ++p) {
                    const std::vector<int>& L = lists[p];
                    f.scp.push_back(L.size() > 0 ? L[0] : -1);
                    if (L.size() <= 2) {
                        f.scp.push_back(L.size() > 1 ? L[1] : -1);
                    } else {
                        f.scp.push_back(-2 - (int)f.sfc.size());
                        f.sfc.insert(f.sfc.end(), L.begin() + 1, L.end());
                        f.sfc.push_back(-1);
                    }
                }
                for (int e = f.roff[t]; e < f.roff[t + 1]; ++e) {
                    int row = f.offrows[e];
                    if (row >= col0 && row < col1) {
                        f.sbo.push_back(row - col0);
                        f.sbo.push_back(-1);
                    } else {  // the owner's partials of x at that row
                        const int* ri = f.rowinfo.data() + 4 * (size_t)row;
                        f.sbo.push_back(ri[0]);
                        f.sbo.push_back((ri[2] << 16) | ri[1]);
                    }
                }
            }
        }
    }
}

// ---- level pairs ------------------------------------------------------------------------

std::vector<int> pair_levels(const std::vector<int>& order, const std::vector<int>& ptr, const std::vector<int>& adj,
                             int sub_budget, double amalg_ratio, std::vector<int>& group) {
    const int m = (int)order.size();
    group.assign(m, -1);
    if (m == 0) return order;
    // the pattern in this order, and its factor as it would be laid out
    std::vector<int> iorder(m);
    for (int k = 0; k < m; ++k) iorder[order[k]] = k;
    std::vector<int> p2(m + 1, 0), a2;
    a2.reserve(adj.size());
    for (int k = 0; k < m; ++k) {
        const int i = order[k];
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) a2.push_back(iorder[adj[e]]);
        p2[k + 1] = (int)a2.size();
    }
    Factor f;
    symbolic(m, p2, a2, f, sub_budget, amalg_ratio);
    const int ns = f.ns, top = f.nlevels - 1;
    if (top < 2) return order;
    // members: level supernodes whose level is an odd distance below the top level and whose
    // parent sits one level up; they join their parent's group
    std::vector<int> head(ns, -1);
    for (int t = ns - 1; t >= 0; --t) {
        if (f.level[t] < 0) continue;
        const int p = f.parent[t];
        const bool member = p >= 0 && (top - f.level[t]) % 2 == 0 && f.level[p] == f.level[t] + 1 && head[p] == p;
        head[t] = member ? p : t;
    }
    // children lists in id order
    std::vector<std::vector<int>> kids(ns);
    std::vector<int> roots;
    for (int t = 0; t < ns; ++t) {
        if (f.parent[t] >= 0) kids[f.parent[t]].push_back(t);
        else roots.push_back(t);
    }
    // new order of positions: a head's members' subtrees first, then the members' columns,
    // then the head's; every column still follows its descendants (same fill)
    std::vector<int> pos;
    pos.reserve(m);
    std::vector<int> gpos;
    gpos.reserve(m);
    auto cols = [&](int t, int g) {
        for (int j = f.c0[t]; j < f.c0[t] + f.w[t]; ++j) { pos.push_back(j); gpos.push_back(g); }
    };
    struct Frame { int t, stage; };
    std::vector<Frame> st;
    for (int r : roots) {
        st.push_back({r, 0});
        while (!st.empty()) {
            const int t = st.back().t;
            const int stage = st.back().stage;
            const std::vector<int>& ch = kids[t];
            if (stage < (int)ch.size()) {
                st.back().stage = stage + 1;
                const int c = ch[stage];
                if (head[c] == t && c != t) {  // a member: its children's subtrees now
                    for (int q = (int)kids[c].size() - 1; q >= 0; --q) st.push_back({kids[c][q], 0});
                } else {
                    st.push_back({c, 0});
                }
                continue;
            }
            const bool is_head = head[t] == t;
            bool grouped = false;
            for (int c : ch)
                if (head[c] == t && c != t) { cols(c, t); grouped = true; }
            cols(t, is_head && grouped ? t : -1);
            st.pop_back();
        }
    }
    std::vector<int> out(m);
    for (int k = 0; k < m; ++k) out[k] = order[pos[k]];
    // members' columns carry the group too
    for (int k = 0; k < m; ++k) group[k] = gpos[k];
    return out;
}

// ---- numeric ----------------------------------------------------------------------------

bool numeric(const std::vector<int>& ptr, const std::vector<int>& adj, const std::vector<double>& val,
             const std::vector<double>& diag, Factor& f, std::string& err) {
    int n = f.n;
    f.panel.assign(f.poff[f.ns], 0.0);
    double amax = 0.0;
    for (int i = 0; i < n; ++i) amax = std::max(amax, std::fabs(diag[i]));
    std::vector<std::vector<double>> upd(f.ns);
    std::vector<int> pos(n, -1);
    std::vector<double> front;
    for (int s = 0; s < f.ns; ++s) {
        const int c0 = f.c0[s], w = f.w[s], no = f.noff[s], m = w + no;
        const int* rows = f.offrows.data() + f.roff[s];
        for (int k = 0; k < w; ++k) pos[c0 + k] = k;
        for (int k = 0; k < no; ++k) pos[rows[k]] = w + k;
        front.assign((size_t)m * m, 0.0);  // column-major, lower part used
        auto F = [&](int r, int c) -> double& { return front[(size_t)c * m + r]; };
        for (int k = 0; k < w; ++k) {
            int j = c0 + k;
            F(k, k) += diag[j];
            for (int e = ptr[j]; e < ptr[j + 1]; ++e) {
                int i = adj[e];
                if (i > j) F(pos[i], k) += val[e];
            }
        }
        for (int q = f.child_ptr[s]; q < f.child_ptr[s + 1]; ++q) {
            int c = f.child[q];
            int nc = f.noff[c];
            const int* rl = f.rel.data() + f.roff[c];
            const std::vector<double>& U = upd[c];
            for (int b = 0; b < nc; ++b)
                for (int a = b; a < nc; ++a) {
                    int ra = rl[a], rb = rl[b];
                    // (ra >= rb since rel is increasing)
                    F(ra, rb) += U[(size_t)b * nc + a];
                }
            std::vector<double>().swap(upd[c]);
        }
        // Cholesky of the pivot block (left-looking, lower)
        std::vector<double> Ls((size_t)w * w, 0.0);
        for (int c = 0; c < w; ++c)
            for (int r = c; r < w; ++r) Ls[(size_t)c * w + r] = F(r, c);
        for (int c = 0; c < w; ++c) {
            for (int k = 0; k < c; ++k) {
                double lck = Ls[(size_t)k * w + c];
                if (lck == 0.0) continue;
                for (int r = c; r < w; ++r) Ls[(size_t)c * w + r] -= Ls[(size_t)k * w + r] * lck;
            }
            double d = Ls[(size_t)c * w + c];
            if (!(d > 1e-14 * amax)) {
                err = "factorization failed: pivot block of supernode " + std::to_string(s) +
                      " is not positive definite (a mesh component without constraints?)";
                return false;
            }
            d = std::sqrt(d);
            Ls[(size_t)c * w + c] = d;
            for (int r = c + 1; r < w; ++r) Ls[(size_t)c * w + r] /= d;
        }
        // Linv = Ls^{-1} (lower), column by column: Ls * Linv[:,c] = e_c
        std::vector<double> Li((size_t)w * w, 0.0);
        for (int c = 0; c < w; ++c) {
            double* x = &Li[(size_t)c * w];
            x[c] = 1.0 / Ls[(size_t)c * w + c];
            for (int k = c; k < w; ++k) {
                double xk = (k == c) ? x[c] : x[k] / Ls[(size_t)k * w + k];
                x[k] = xk;
                for (int r = k + 1; r < w; ++r) x[r] -= Ls[(size_t)k * w + r] * xk;
            }
        }
        // Dinv = Linv^T Linv
        double* P = f.panel.data() + f.poff[s];
        for (int c = 0; c < w; ++c)
            for (int r = c; r < w; ++r) {
                double acc = 0.0;
                const double* lr = &Li[(size_t)r * w];
                const double* lc = &Li[(size_t)c * w];
                for (int k = r; k < w; ++k) acc += lr[k] * lc[k];
                P[(size_t)c * m + r] = acc;
                P[(size_t)r * m + c] = acc;
            }
        if (no > 0) {
            // G = F_os * Dinv  (no x w), stored in panel rows [w, m)
            for (int q = 0; q < w; ++q) {
                double* g = P + (size_t)q * m + w;
                for (int k = 0; k < w; ++k) {
                    double dkq = P[(size_t)q * m + k];
                    const double* fo = &front[(size_t)k * m + w];
                    for (int r = 0; r < no; ++r) g[r] += fo[r] * dkq;
                }
            }
            // U = F_oo - G F_os^T (lower part, column-major no x no)
            std::vector<double>& U = upd[s];
            U.assign((size_t)no * no, 0.0);
            for (int b = 0; b < no; ++b)
                for (int a = b; a < no; ++a) U[(size_t)b * no + a] = F(w + a, w + b);
            for (int k = 0; k < w; ++k) {
                const double* g = P + (size_t)k * m + w;
                const double* fo = &front[(size_t)k * m + w];
                for (int b = 0; b < no; ++b) {
                    double fb = fo[b];
                    if (fb == 0.0) continue;
                    double* u = &U[(size_t)b * no];
                    for (int a = b; a < no; ++a) u[a] -= g[a] * fb;
                }
            }
        }
        for (int k = 0; k < w; ++k) pos[c0 + k] = -1;
        for (int k = 0; k < no; ++k) pos[rows[k]] = -1;
    }
    return true;
}

// The device numeric factorization run on the host, step for step as the kernels run it
// (one height at a time; per supernode: zeroed workspace, original entries, children's update
// matrices in child order, left-looking Cholesky of the pivot block, its inverse, D^{-1},
// G = F_os D^{-1}, the update matrix), with every index checked against its block. Leaves the
// FP64 panels in f.panel.
bool numeric_device_emulation(const std::vector<int>& ptr, const std::vector<int>& adj,
                              const std::vector<double>& val, const std::vector<double>& diag, Factor& f,
                              std::string& err) {
    const int n = f.n, ns = f.ns;
    std::vector<double> P(f.poff[ns], 0.0), S(f.woff_s[ns], 0.0), T(f.woff_t[ns], 0.0), U(f.woff_u[ns], 0.0);
    double amax = 0.0;
    for (int i = 0; i < n; ++i) amax = std::max(amax, std::fabs(diag[i]));
#define MD_FCHECK(cond, what)                                  \
    if (!(cond)) {                                             \
        err = std::string("factor emulation: ") + what;        \
        return false;                                          \
    }
    const int nh = (int)f.hptr.size() - 1;
    for (int h = 0; h < nh; ++h)
        for (int hi = f.hptr[h]; hi < f.hptr[h + 1]; ++hi) {
            const int t = f.hlist[hi];
            const int c0 = f.c0[t], w = f.w[t], no = f.noff[t], m = w + no;
            const int64_t po = f.poff[t], so = f.woff_s[t], to = f.woff_t[t], uo = f.woff_u[t];
            auto Pp = [&](int r, int c) -> double& { return P[po + (int64_t)c * m + r]; };
            auto Sp = [&](int r, int c) -> double& { return S[so + (int64_t)c * no + r]; };
            auto Tp = [&](int r, int c) -> double& { return T[to + (int64_t)r * w + c]; };  // row-major
            auto Up = [&](int r, int c) -> double& { return U[uo + (int64_t)c * no + r]; };
            const int* rows = f.offrows.data() + f.roff[t];
            // original entries, one column at a time
            for (int k = 0; k < w; ++k) {
                const int j = c0 + k;
                MD_FCHECK(f.col_sn[j] == t, "column owner");
                Pp(k, k) = diag[j];
                for (int e = ptr[j]; e < ptr[j + 1]; ++e) {
                    const int i = adj[e];
                    if (i <= j) continue;
                    if (i < c0 + w) {
                        Pp(i - c0, k) += val[e];
                    } else {
                        const int r = (int)(std::lower_bound(rows, rows + no, i) - rows);
                        MD_FCHECK(r < no && rows[r] == i, "off row of an original entry");
                        Sp(r, k) += val[e];
                    }
                }
            }
            // children's update matrices, in child order
            for (int q = f.child_ptr[t]; q < f.child_ptr[t + 1]; ++q) {
                const int c = f.child[q], nc = f.noff[c];
                const int* rel = f.rel.data() + f.roff[c];
                for (int64_t idx = 0; idx < (int64_t)nc * nc; ++idx) {
                    const int b = (int)(idx / nc), a = (int)(idx % nc);
                    if (a < b) continue;
                    const double v = U[f.woff_u[c] + (int64_t)b * nc + a];
                    const int ra = rel[a], rb = rel[b];
                    MD_FCHECK(ra >= rb && ra < m && rb >= 0, "update position");
                    if (rb < w) {
                        if (ra < w) Pp(ra, rb) += v;
                        else Sp(ra - w, rb) += v;
                    } else {
                        Up(ra - w, rb - w) += v;
                    }
                }
            }
            // Cholesky factor and its inverse X of the pivot block, blocked by 64 columns
            // (right-looking), in the device's workspaces; the device orders the work within a
            // front differently (an unscaled LDL^T chain per block, the X blocks inside the
            // right-looking loop, products split over K), so the two agree to rounding.
            constexpr int kFb = 64;
            const int nb = (w + kFb - 1) / kFb;
            for (int s = 0; s < nb; ++s) {
                const int kb = s * kFb, b = std::min(kFb, w - kb);
                // diagonal block: unblocked Cholesky, then its inverse (row-major, upper part zero)
                for (int k = 0; k < b; ++k) {
                    const double d0 = Pp(kb + k, kb + k);
                    if (!(d0 > 1e-14 * amax)) {
                        err = "factorization failed: pivot block of supernode " + std::to_string(t) +
                              " is not positive definite (a mesh component without constraints?)";
                        return false;
                    }
                    const double d = std::sqrt(d0);
                    for (int r = k + 1; r < b; ++r) Pp(kb + r, kb + k) /= d;
                    Pp(kb + k, kb + k) = d;
                    for (int c = k + 1; c < b; ++c)
                        for (int r = c; r < b; ++r) Pp(kb + r, kb + c) -= Pp(kb + r, kb + k) * Pp(kb + c, kb + k);
                }
                for (int r = 0; r < b; ++r)
                    for (int c = 0; c < b; ++c) Tp(kb + r, kb + c) = r == c ? 1.0 : 0.0;
                for (int k = 0; k < b; ++k) {
                    const double lkk = Pp(kb + k, kb + k);
                    for (int c = 0; c <= k; ++c) Tp(kb + k, kb + c) = c == k ? 1.0 / lkk : Tp(kb + k, kb + c) / lkk;
                    for (int r = k + 1; r < b; ++r)
                        for (int c = 0; c <= k; ++c) Tp(kb + r, kb + c) -= Pp(kb + r, kb + k) * Tp(kb + k, kb + c);
                }
                if (s == nb - 1) break;
                // rows below: forward substitution against the diagonal block
                for (int r = kb + kFb; r < w; ++r)
                    for (int j = 0; j < kFb; ++j) {
                        double v = Pp(r, kb + j);
                        for (int q = 0; q < j; ++q) v -= Pp(r, kb + q) * Pp(kb + j, kb + q);
                        Pp(r, kb + j) = v / Pp(kb + j, kb + j);
                    }
                // trailing lower part: A(r, c) -= sum_q L(r, kb + q) L(c, kb + q)
                for (int c = kb + kFb; c < w; ++c)
                    for (int r = c; r < w; ++r) {
                        double acc = 0.0;
                        for (int q = 0; q < kFb; ++q) acc += Pp(r, kb + q) * Pp(c, kb + q);
                        Pp(r, c) -= acc;
                    }
            }
            // off-diagonal blocks of X, one block diagonal at a time:
            // X_IJ = -X_II sum_{K=J}^{I-1} L_IK X_KJ
            for (int d = 1; d < nb; ++d)
                for (int I = d; I < nb; ++I) {
                    const int J = I - d, r0 = I * kFb, c0 = J * kFb, rb = std::min(kFb, w - r0);
                    std::vector<double> Y((size_t)kFb * kFb, 0.0);
                    for (int K = J; K < I; ++K)
                        for (int k = 0; k < kFb; ++k)
                            for (int x = 0; x < rb; ++x)
                                for (int y = 0; y < kFb; ++y)
                                    Y[(size_t)x * kFb + y] += Pp(r0 + x, K * kFb + k) * Tp(K * kFb + k, c0 + y);
                    for (int x = 0; x < rb; ++x)
                        for (int y = 0; y < kFb; ++y) {
                            double acc = 0.0;
                            for (int k = 0; k < rb; ++k) acc += Tp(r0 + x, r0 + k) * Y[(size_t)k * kFb + y];
                            Tp(r0 + x, c0 + y) = -acc;
                        }
                }
            // D^{-1} = Linv^T Linv (symmetric)
            for (int r = 0; r < w; ++r)
                for (int c = 0; c <= r; ++c) {
                    double acc = 0.0;
                    for (int k = r; k < w; ++k) acc += Tp(k, r) * Tp(k, c);
                    Pp(r, c) = acc;
                    Pp(c, r) = acc;
                }
            if (no > 0) {
                // G = F_os D^{-1} (panel rows [w, m))
                for (int q = 0; q < w; ++q)
                    for (int r = 0; r < no; ++r) {
                        double acc = 0.0;
                        for (int k = 0; k < w; ++k) acc += Sp(r, k) * Pp(k, q);
                        Pp(w + r, q) = acc;
                    }
                // U -= G F_os^T (lower)
                for (int b = 0; b < no; ++b)
                    for (int a = b; a < no; ++a) {
                        double acc = Up(a, b);
                        for (int k = 0; k < w; ++k) {
                            const double fb = Sp(b, k);
                            if (fb != 0.0) acc -= Pp(w + a, k) * fb;
                        }
                        Up(a, b) = acc;
                    }
            }
        }
#undef MD_FCHECK
    f.panel.swap(P);
    return true;
}

void solve_host(const Factor& f, const double* b, double* x, int nrhs) {
    int n = f.n;
    for (int r = 0; r < nrhs; ++r) {
        const double* br = b + (size_t)r * n;
        double* xr = x + (size_t)r * n;
        std::vector<double> z(n), u(f.roff[f.ns]);
        for (int s = 0; s < f.ns; ++s) {
            int c0 = f.c0[s], w = f.w[s], no = f.noff[s], m = w + no;
            std::vector<double> v(m, 0.0);
            for (int k = 0; k < w; ++k) v[k] = br[c0 + k];
            for (int q = f.child_ptr[s]; q < f.child_ptr[s + 1]; ++q) {
                int c = f.child[q];
                for (int k = 0; k < f.noff[c]; ++k) v[f.rel[f.roff[c] + k]] += u[f.roff[c] + k];
            }
            const double* P = f.panel.data() + f.poff[s];
            for (int k = 0; k < w; ++k) z[c0 + k] = v[k];
            for (int rr = 0; rr < no; ++rr) {
                double acc = v[w + rr];
                for (int q = 0; q < w; ++q) acc -= P[(size_t)q * m + w + rr] * v[q];
                u[f.roff[s] + rr] = acc;
            }
        }
        for (int s = f.ns - 1; s >= 0; --s) {
            int c0 = f.c0[s], w = f.w[s], no = f.noff[s], m = w + no;
            const int* rows = f.offrows.data() + f.roff[s];
            const double* P = f.panel.data() + f.poff[s];
            for (int q = 0; q < w; ++q) {
                double acc = 0.0;
                for (int p = 0; p < w; ++p) acc += P[(size_t)q * m + p] * z[c0 + p];
                for (int rr = 0; rr < no; ++rr) acc -= P[(size_t)q * m + w + rr] * xr[rows[rr]];
                xr[c0 + q] = acc;
            }
        }
    }
}

// Emulates the device sweeps tile by tile (same tile lists, bucket pointers, partial
// vectors and locators as k_fwd / k_bwd / k_apply), checking every index against the
// buffer it reads or writes. Returns false on an out-of-range access.
bool solve_tiles_host(const Factor& f, const double* b, double* x, std::string& err) {
    const int n = f.n;
    const int64_t nu = f.ubase.empty() ? 0 : f.ubase.back(), nx = f.xoff.empty() ? 0 : f.xoff.back();
    std::vector<double> z(3 * (size_t)n, 0.0), up(3 * (size_t)nu, 0.0), xp(3 * (size_t)nx, 0.0);
    std::vector<double> pc(f.pc_off.back()), pr(f.pr_off.back());
    for (int s = 0; s < f.ns; ++s) {
        int w = f.w[s], no = f.noff[s], mm = w + no;
        const double* P = f.panel.data() + f.poff[s];
        for (int q = 0; q < w; ++q) {
            for (int r = 0; r < no; ++r) pc[f.pc_off[s] + (size_t)q * no + r] = P[(size_t)q * mm + w + r];
            for (int p = 0; p < mm; ++p) pr[f.pr_off[s] + (size_t)p * w + q] = P[(size_t)q * mm + p];
        }
    }
#define MD_CHECK(cond, what)                       \
    if (!(cond)) {                                 \
        err = std::string("tile emulation: ") + what; \
        return false;                              \
    }
    auto sum_part = [&](const std::vector<double>& buf, int64_t base, int cnt, int stride, int64_t e, int c,
                        double& out) -> bool {
        out = 0.0;
        for (int k = 0; k < cnt; ++k) {
            int64_t idx = base + (int64_t)k * stride + e;
            if (idx < 0 || 3 * idx + c >= (int64_t)buf.size()) return false;
            out += buf[3 * idx + c];
        }
        return true;
    };
    // bottom subtrees, as the kernels run them: records of the subtree, lanes mapped to
    // (supernode, item) by binary search over the prefixes of a height. LDS entries carry
    // the phase that wrote them; a read must see an entry written in an earlier phase (a
    // same-phase read would race on the device).
    std::vector<double> lds;
    std::vector<int> stamp;
    auto lds_read = [&](int e, int entries, int phase, int c, double& v) -> bool {
        if (e < 0 || e >= entries || stamp[e] <= 0 || stamp[e] >= phase) return false;
        v = lds[3 * (size_t)e + c];
        return true;
    };
    auto lds_claim = [&](int e, int entries, int phase) -> bool {
        if (e < 0 || e >= entries || stamp[e] == phase) return false;
        stamp[e] = phase;
        return true;
    };
    // records [a, b) of one height; field: 11 positions, 12 off rows, 13 columns; count of
    // the items of record j: m, noff or w
    auto rec = [&](int j) { return f.snode.data() + 16 * (size_t)j; };
    auto items = [&](int j, int field) {
        const int* R = rec(j);
        return field == 11 ? R[1] + R[2] : field == 12 ? R[2] : R[1];
    };
    auto find = [&](int a, int b, int field, int t) {
        int lo = a, hi = b - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (rec(mid)[field] <= t) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    };
    auto sub_heights = [&](int i, int h, int& a, int& b) -> bool {
        size_t e = (size_t)f.sub[4 * (size_t)i + 2] + h;
        if (e + 1 >= f.sub_h.size()) return false;
        a = f.sub_h[e];
        b = f.sub_h[e + 1];
        return a >= f.sub[4 * (size_t)i] && a < b && b <= f.sub[4 * (size_t)i + 1] && 16 * (size_t)b <= f.snode.size();
    };
    // LDS entries one subtree uses in each sweep
    auto sub_entries = [&](int i, bool fwd) {
        int e = 0;
        for (int j = f.sub[4 * (size_t)i]; j < f.sub[4 * (size_t)i + 1]; ++j) {
            const int* R = rec(j);
            if (fwd) e = std::max(e, R[3] + R[1] + R[2]);
            else e = std::max(e, std::max(R[10] + R[1], R[14] + R[1] + R[2]));
        }
        return e;
    };
    MD_CHECK((int)f.sub.size() == 4 * f.nsub, "subtree table");
    for (int i = 0; i < f.nsub; ++i) {
        const int nh = f.sub[4 * (size_t)i + 3], entries = sub_entries(i, true);
        MD_CHECK(entries > 0 && entries <= f.sub_lds_fwd && nh > 0, "subtree extent");
        MD_CHECK(f.sub[4 * (size_t)i + 1] - f.sub[4 * (size_t)i] <= f.sub_nodes, "subtree records");
        lds.assign(3 * (size_t)entries, 0.0);
        stamp.assign(entries, 0);
        int phase = 1;
        // the one-mesh form first writes the residual of the subtree's rows into their front
        // positions (sfi); the gathers below read it from there
        MD_CHECK(2 * (size_t)i + 1 < f.srow.size(), "subtree row range");
        const int col0 = f.srow[2 * (size_t)i], ncols = f.srow[2 * (size_t)i + 1];
        MD_CHECK(col0 >= 0 && ncols > 0 && (size_t)col0 + ncols <= f.sfi.size(), "subtree rows");
        for (int t = 0; t < ncols; ++t) {
            const int e = f.sfi[(size_t)col0 + t];
            MD_CHECK(lds_claim(e, entries, phase), "subtree residual write");
            for (int c = 0; c < 3; ++c) lds[3 * (size_t)e + c] = b[(size_t)c * n + col0 + t];
        }
        for (int h = 0; h < nh; ++h) {
            int ra, rb;
            MD_CHECK(sub_heights(i, h, ra, rb), "subtree height range");
            ++phase;
            const int npos = rec(rb - 1)[11] + items(rb - 1, 11);
            for (int t = 0; t < npos; ++t) {
                const int j = find(ra, rb, 11, t);
                const int* R = rec(j);
                const int p = t - R[11], w = R[1];
                MD_CHECK(p >= 0 && p < w + R[2], "subtree position map");
                MD_CHECK(R[6] >= 0 && 2 * ((size_t)R[6] + p) + 1 < f.scp.size(), "subtree contribution entry");
                // child entries: the first two inline, the rest from a -1-terminated list
                std::vector<int> src;
                const int e0 = f.scp[2 * (R[6] + p)], e1 = f.scp[2 * (R[6] + p) + 1];
                if (e0 >= 0) src.push_back(e0);
                if (e1 >= 0) {
                    src.push_back(e1);
                } else if (e1 <= -2) {
                    MD_CHECK(e0 >= 0, "subtree overflow without a first entry");
                    for (size_t c = (size_t)(-2 - e1);; ++c) {
                        MD_CHECK(c < f.sfc.size(), "subtree overflow list");
                        if (f.sfc[c] < 0) break;
                        src.push_back(f.sfc[c]);
                    }
                }
                double v[3] = {0, 0, 0};
                for (int c = 0; c < 3; ++c) {
                    if (p < w) {
                        MD_CHECK(R[0] + p < n && f.sfi[R[0] + p] == R[3] + p, "subtree residual slot");
                        MD_CHECK(lds_read(R[3] + p, entries, phase, c, v[c]), "subtree residual read");
                        MD_CHECK(v[c] == b[(size_t)c * n + R[0] + p], "subtree residual value");
                    }
                    for (int e : src) {
                        double u;
                        MD_CHECK(lds_read(e, entries, phase, c, u), "subtree child update read");
                        v[c] += u;
                    }
                }
                MD_CHECK(lds_claim(R[3] + p, entries, phase), "subtree front write");
                for (int c = 0; c < 3; ++c) {
                    lds[3 * (size_t)(R[3] + p) + c] = v[c];
                    if (p < w) z[3 * (size_t)(R[0] + p) + c] = v[c];
                }
            }
            ++phase;
            const int nrow = rec(rb - 1)[12] + items(rb - 1, 12);
            for (int t = 0; t < nrow; ++t) {
                const int j = find(ra, rb, 12, t);
                const int* R = rec(j);
                const int i_ = t - R[12], w = R[1], no = R[2];
                MD_CHECK(i_ >= 0 && i_ < no && w <= kTile, "subtree off-row map");
                MD_CHECK(R[4] >= 0 && (int64_t)R[4] + i_ + (int64_t)(w - 1) * no < (int64_t)pc.size(), "subtree G range");
                double acc[3] = {0, 0, 0}, o[3];
                for (int q = 0; q < w; ++q)
                    for (int c = 0; c < 3; ++c) {
                        double v;
                        MD_CHECK(lds_read(R[3] + q, entries, phase, c, v), "subtree front read");
                        acc[c] += pc[R[4] + i_ + (size_t)q * no] * v;
                    }
                const int dst = R[3] + w + i_;
                for (int c = 0; c < 3; ++c) MD_CHECK(lds_read(dst, entries, phase, c, o[c]), "subtree off-row read");
                MD_CHECK(lds_claim(dst, entries, phase), "subtree update write");
                int64_t udst = -1;
                if (R[8] >= 0) {  // the root: its update entry goes to the parent's slot
                    MD_CHECK((size_t)R[8] + i_ < f.udst.size(), "subtree update entry");
                    udst = f.udst[R[8] + i_];
                    MD_CHECK(udst >= 0 && udst < nu, "subtree update output");
                }
                for (int c = 0; c < 3; ++c) {
                    double u = o[c] - acc[c];
                    lds[3 * (size_t)dst + c] = u;
                    if (udst >= 0) up[3 * (size_t)udst + c] = u;
                }
            }
        }
    }
    // children's contributions to position pos of level supernode s: the sum of its uk slots
    // (unused ones zero), in slot order as the kernels sum them
    auto pull_at = [&](int s, int pos, double* acc3) -> bool {
        const int m = f.w[s] + f.noff[s];
        if (pos < 0 || pos >= m || f.uk[s] < 0) return false;
        if (f.ubase[s] + (int64_t)f.uk[s] * m > f.ubase[s + 1]) return false;
        for (int c = 0; c < 3; ++c) {
            double v;
            if (!sum_part(up, f.ubase[s], f.uk[s], m, pos, c, v)) return false;
            acc3[c] += v;
        }
        return true;
    };
    for (int l = 0; l < f.nlevels; ++l) {
        for (int t = f.ft_ptr[l]; t < f.ft_ptr[l + 1]; ++t) {
            int s = f.ft[4 * (size_t)t], r0 = f.ft[4 * (size_t)t + 1], q0 = f.ft[4 * (size_t)t + 2];
            MD_CHECK(s >= 0 && s < f.ns && f.level[s] == l, "forward tile supernode");
            int w = f.w[s], no = f.noff[s], c0 = f.c0[s];
            int qn = std::min(kTile, w - q0), rn = std::max(0, std::min(kTile, no - r0)), kcb = q0 / kTile;
            MD_CHECK(qn > 0 && q0 % kTile == 0 && r0 % kTile == 0, "forward tile origin");
            double vd[3][kTile] = {}, vo[3][kTile] = {};
            for (int i = 0; i < qn; ++i)
                for (int c = 0; c < 3; ++c) vd[c][i] = b[(size_t)c * n + c0 + q0 + i];
            auto pull = [&](int pos, double* acc3) -> bool { return pull_at(s, pos, acc3); };
            for (int i = 0; i < qn; ++i) {
                double a[3] = {0, 0, 0};
                MD_CHECK(pull(q0 + i, a), "column contributions");
                for (int c = 0; c < 3; ++c) vd[c][i] += a[c];
            }
            if (kcb == 0)
                for (int i = 0; i < rn; ++i) {
                    double a[3] = {0, 0, 0};
                    MD_CHECK(pull(w + r0 + i, a), "row contributions");
                    for (int c = 0; c < 3; ++c) vo[c][i] += a[c];
                }
            if (r0 == 0)
                for (int i = 0; i < qn; ++i)
                    for (int c = 0; c < 3; ++c) z[3 * (size_t)(c0 + q0 + i) + c] = vd[c][i];
            for (int lane = 0; lane < rn; ++lane) {
                double acc[3] = {0, 0, 0};
                for (int i = 0; i < qn; ++i) {
                    int64_t idx = f.pc_off[s] + (int64_t)(q0 + i) * no + r0 + lane;
                    MD_CHECK(idx < f.pc_off[s + 1], "forward panel");
                    for (int c = 0; c < 3; ++c) acc[c] += pc[idx] * vd[c][i];
                }
                const int par = f.parent[s];
                MD_CHECK(par >= 0 && (size_t)f.roff[s] + r0 + lane < f.udst.size(), "update destination");
                const int64_t o = f.udst[f.roff[s] + r0 + lane] + (int64_t)kcb * (f.w[par] + f.noff[par]);
                MD_CHECK(f.udst[f.roff[s] + r0 + lane] >= f.ubase[par] && o < f.ubase[par + 1], "update output");
                for (int c = 0; c < 3; ++c) up[3 * o + c] = (kcb == 0 ? vo[c][lane] : 0.0) - acc[c];
            }
        }
    }
    for (int l = f.nlevels - 1; l >= 0; --l) {
        for (int t = f.bt_ptr[l]; t < f.bt_ptr[l + 1]; ++t) {
            int s = f.bt[4 * (size_t)t], q0 = f.bt[4 * (size_t)t + 1], p0 = f.bt[4 * (size_t)t + 2];
            MD_CHECK(s >= 0 && s < f.ns && f.level[s] == l, "backward tile supernode");
            int w = f.w[s], mm = w + f.noff[s];
            int qn = std::min(kTile, w - q0), pn = std::min(kTile, mm - p0);
            MD_CHECK(qn > 0 && pn > 0, "backward tile extent");
            double tv[3][kTile] = {};
            for (int lane = 0; lane < pn; ++lane) {
                int p = p0 + lane;
                for (int c = 0; c < 3; ++c) {
                    if (p < w && f.noff[s] == 0) {
                        tv[c][lane] = b[(size_t)c * n + f.c0[s] + p];  // a root: no forward tile
                    } else if (p < w) {
                        tv[c][lane] = z[3 * (size_t)(f.c0[s] + p) + c];
                    } else {
                        size_t e = (size_t)f.roff[s] + p - w;
                        MD_CHECK(e < f.offrows.size(), "off row entry");
                        const int* xl = f.xloc.data() + 4 * e;
                        double v;
                        MD_CHECK(sum_part(xp, xl[0], xl[2], xl[1], 0, c, v), "solution partial");
                        tv[c][lane] = -v;
                    }
                }
                if (p < w && f.noff[s] == 0) {
                    double a[3] = {0, 0, 0};
                    MD_CHECK(pull_at(s, p, a), "root contributions");
                    for (int c = 0; c < 3; ++c) tv[c][lane] += a[c];
                }
            }
            for (int lane = 0; lane < qn; ++lane) {
                double acc[3] = {0, 0, 0};
                for (int i = 0; i < pn; ++i) {
                    int64_t idx = f.pr_off[s] + (int64_t)(p0 + i) * w + q0 + lane;
                    MD_CHECK(idx < f.pr_off[s + 1], "backward panel");
                    for (int c = 0; c < 3; ++c) acc[c] += pr[idx] * tv[c][i];
                }
                int64_t o = f.xoff[s] + (int64_t)(p0 / kTile) * w + q0 + lane;
                MD_CHECK(o < f.xoff[s + 1], "solution output");
                for (int c = 0; c < 3; ++c) xp[3 * o + c] = acc[c];
            }
        }
    }
    for (int i = 0; i < f.nsub; ++i) {
        const int nh = f.sub[4 * (size_t)i + 3], entries = sub_entries(i, false);
        MD_CHECK(entries > 0 && entries <= f.sub_lds_bwd, "subtree extent (backward)");
        lds.assign(3 * (size_t)entries, 0.0);
        stamp.assign(entries, 0);
        // the column region starts with z of the subtree's rows; each row's x replaces it
        int phase = 1;
        const int col0 = f.srow[2 * (size_t)i], ncols = f.srow[2 * (size_t)i + 1];
        for (int t = 0; t < ncols; ++t) {
            MD_CHECK(lds_claim(t, entries, phase), "subtree z write");
            for (int c = 0; c < 3; ++c) lds[3 * (size_t)t + c] = z[3 * (size_t)(col0 + t) + c];
        }
        for (int h = nh - 1; h >= 0; --h) {
            int ra, rb;
            MD_CHECK(sub_heights(i, h, ra, rb), "subtree height range");
            ++phase;
            const int npos = rec(rb - 1)[11] + items(rb - 1, 11);
            for (int t = 0; t < npos; ++t) {
                const int j = find(ra, rb, 11, t);
                const int* R = rec(j);
                const int p = t - R[11], w = R[1];
                double v[3];
                if (p < w) {
                    MD_CHECK(R[10] + p < ncols && col0 + R[10] + p == R[0] + p, "subtree column slot");
                    for (int c = 0; c < 3; ++c) MD_CHECK(lds_read(R[10] + p, entries, phase, c, v[c]), "subtree z read");
                } else {
                    const size_t e = (size_t)R[7] + p - w;
                    MD_CHECK(R[7] >= 0 && 2 * e + 1 < f.sbo.size(), "subtree off-row source");
                    const int src = f.sbo[2 * e], info = f.sbo[2 * e + 1];
                    for (int c = 0; c < 3; ++c) {
                        if (info < 0) {
                            MD_CHECK(lds_read(src, entries, phase, c, v[c]), "subtree solution read");
                        } else {
                            MD_CHECK((info >> 16) >= 1, "subtree outer partial count");
                            MD_CHECK(sum_part(xp, src, info >> 16, info & 0xffff, 0, c, v[c]), "subtree outer solution");
                        }
                        v[c] = -v[c];
                    }
                }
                MD_CHECK(lds_claim(R[14] + p, entries, phase), "subtree rhs write");
                for (int c = 0; c < 3; ++c) lds[3 * (size_t)(R[14] + p) + c] = v[c];
            }
            ++phase;
            const int ncol = rec(rb - 1)[13] + items(rb - 1, 13);
            for (int t = 0; t < ncol; ++t) {
                const int j = find(ra, rb, 13, t);
                const int* R = rec(j);
                const int q = t - R[13], w = R[1], mm = w + R[2];
                MD_CHECK(q >= 0 && q < w && mm <= kSubMaxRows && R[5] >= 0 &&
                         (int64_t)R[5] + q + (int64_t)(mm - 1) * w < (int64_t)pr.size(), "subtree P range");
                double acc[3] = {0, 0, 0};
                for (int p = 0; p < mm; ++p)
                    for (int c = 0; c < 3; ++c) {
                        double v;
                        MD_CHECK(lds_read(R[14] + p, entries, phase, c, v), "subtree rhs read");
                        acc[c] += pr[R[5] + q + (size_t)p * w] * v;
                    }
                MD_CHECK(lds_claim(R[10] + q, entries, phase), "subtree solution write");
                MD_CHECK(R[9] >= 0 && R[9] + q < nx, "subtree solution output");
                for (int c = 0; c < 3; ++c) {
                    lds[3 * (size_t)(R[10] + q) + c] = acc[c];
                    xp[3 * (size_t)(R[9] + q) + c] = acc[c];
                }
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        const int* ri = f.rowinfo.data() + 4 * (size_t)i;
        for (int c = 0; c < 3; ++c) {
            double v;
            MD_CHECK(sum_part(xp, ri[0], ri[2], ri[1], 0, c, v), "final partial");
            x[(size_t)c * n + i] = v;
        }
    }
#undef MD_CHECK
    return true;
}

AnalysisStats stats(const Topology& t, const Factor& f) {
    AnalysisStats s{};
    s.n = t.n;
    s.nf = t.nf;
    s.nslices = t.nslices;
    s.ns = f.ns;
    s.nlevels = f.nlevels;
    s.slots = t.slice_ptr.empty() ? 0 : t.total_slots();
    s.panel_entries = f.factor_nnz();
    int64_t strict = 0;
    for (int k = 0; k < f.ns; ++k)
        strict += (int64_t)f.w[k] * (f.w[k] - 1) / 2 + (int64_t)f.w[k] * f.noff[k];
    s.strict_nnz_equiv = strict;
    s.nnz_exact = f.nnz_exact;
    s.fwd_items = f.ft_ptr.empty() ? 0 : f.ft_ptr.back();
    s.bwd_items = f.bt_ptr.empty() ? 0 : f.bt_ptr.back();
    s.max_width = 0;
    s.max_rows = 0;
    for (int k = 0; k < f.ns; ++k) {
        s.max_width = std::max(s.max_width, f.w[k]);
        s.max_rows = std::max(s.max_rows, f.w[k] + f.noff[k]);
    }
    return s;
}

}  // namespace md
