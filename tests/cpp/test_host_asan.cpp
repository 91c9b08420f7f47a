// Host analysis, factorization and sweep emulation under AddressSanitizer and
// UndefinedBehaviorSanitizer (built by this directory's Makefile with -fsanitize=address,
// undefined; any report aborts the program with a non-zero status).
//
//   test_host_asan mesh FILE   a mesh with constrained vertices: the mesh analysis of a
//                              dirty deform() (free-vertex ordering with and without level
//                              pairs, one-ring topology, free-free pattern, symbolic factor),
//                              the FP64 factorization (host and device emulation) of the
//                              graph Laplacian with the constrained vertices eliminated,
//                              re-laid out for every subtree budget and solved by the
//                              bounds-checked emulation of the device sweeps; prints the
//                              largest relative residual.
//   test_host_asan system FILE a symmetric positive definite system (full CSR, coordinates,
//                              right-hand sides) through md_host_factor_solve for every
//                              budget, level-pair and factorization setting; prints the
//                              largest relative residual.
//
// FILE layouts (little-endian): mesh: int32 n, F, k; int32 tri[3F]; float64 xyz[3n];
// int32 constrained[k]. system: int32 n, nnz, nrhs; int32 rowptr[n+1], col[nnz];
// float64 val[nnz], xyz[3n], b[nrhs * n].
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "md_analysis.h"
#include "md_arap.h"

namespace md {
void set_error(const std::string &msg) { std::fprintf(stderr, "md error: %s\n", msg.c_str()); }
}  // namespace md

template <class T> static bool get(std::ifstream &f, T *p, size_t n) {
    return (bool)f.read(reinterpret_cast<char *>(p), (std::streamsize)(sizeof(T) * n));
}

static const int kBudgets[] = {0, 64, 128, 256, 512, 1024};

// Relative residual max_r |A x_r - b_r| / max |b_r| of a Laplacian given as diag + adjacency.
static double residual(const std::vector<int> &ptr, const std::vector<int> &adj, const std::vector<double> &av,
                       const std::vector<double> &diag, const std::vector<double> &b, const std::vector<double> &x) {
    const int n = (int)diag.size();
    double worst = 0.0;
    for (int r = 0; r < 3; ++r) {
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) {
            double y = diag[i] * x[(size_t)r * n + i];
            for (int e = ptr[i]; e < ptr[i + 1]; ++e) y += av[e] * x[(size_t)r * n + adj[e]];
            num = std::fmax(num, std::fabs(y - b[(size_t)r * n + i]));
            den = std::fmax(den, std::fabs(b[(size_t)r * n + i]));
        }
        worst = std::fmax(worst, num / (den > 0 ? den : 1.0));
    }
    return worst;
}

static int run_mesh(const char *path) {
    std::ifstream f(path, std::ios::binary);
    int32_t h[3];
    if (!get(f, h, 3)) return 2;
    const int n = h[0], F = h[1], k = h[2];
    std::vector<int> tri(3 * (size_t)F), con(k);
    std::vector<double> xyz(3 * (size_t)n);
    if (!get(f, tri.data(), tri.size()) || !get(f, xyz.data(), xyz.size()) || !get(f, con.data(), con.size())) return 2;
    std::vector<char> is_con(n, 0);
    for (int v : con) is_con[v] = 1;
    double worst = 0.0;
    for (int pairs = 0; pairs < 2; ++pairs) {
        std::vector<int> group;
        std::vector<int> perm = md::order_free_vertices(n, F, tri.data(), xyz.data(), is_con, pairs ? &group : nullptr,
                                                        md::kAmalgRatioDefault);
        md::Topology t;
        md::build_topology(n, F, tri.data(), perm, is_con, t);
        const int nf = t.nf;
        // free-free pattern in solve order from the slots; unit weights: L_ii = valence
        std::vector<int> ptr(nf + 1, 0), adj;
        std::vector<double> av, diag(nf, 0.0);
        for (int i = 0; i < nf; ++i) {
            const int s = i / md::kSlice, l = i % md::kSlice;
            for (int q = t.slice_ptr[s]; q < t.slice_ptr[s + 1]; ++q) {
                const int j = t.sell_col[(int64_t)q * md::kSlice + l];
                if (j == i) continue;
                diag[i] += 1.0;
                if (j < nf) { adj.push_back(j); av.push_back(-1.0); }
            }
            ptr[i + 1] = (int)adj.size();
        }
        for (int emu = 0; emu < 2; ++emu) {
            md::Factor fac;
            md::symbolic(nf, ptr, adj, fac, md::kSubBudgetDefault, md::kAmalgRatioDefault, pairs ? &group : nullptr);
            std::string err;
            const bool ok = emu ? md::numeric_device_emulation(ptr, adj, av, diag, fac, err)
                                : md::numeric(ptr, adj, av, diag, fac, err);
            if (!ok) { std::printf("factorization failed: %s\n", err.c_str()); return 1; }
            std::vector<double> b(3 * (size_t)nf), x(3 * (size_t)nf);
            for (size_t i = 0; i < b.size(); ++i) b[i] = std::sin(0.37 * (double)i) + 0.1;
            for (int budget : kBudgets) {
                md::device_layout(fac, budget);
                if (!md::solve_tiles_host(fac, b.data(), x.data(), err)) { std::printf("sweep emulation: %s\n", err.c_str()); return 1; }
                worst = std::fmax(worst, residual(ptr, adj, av, diag, b, x));
            }
            md::AnalysisStats st = md::stats(t, fac);
            if (st.nf != nf) return 1;
        }
    }
    std::printf("mesh residual %.3e\n", worst);
    return 0;
}

static int run_system(const char *path) {
    std::ifstream f(path, std::ios::binary);
    int32_t h[3];
    if (!get(f, h, 3)) return 2;
    const int n = h[0], nnz = h[1], nrhs = h[2];
    std::vector<int> rp(n + 1), col(nnz);
    std::vector<double> val(nnz), xyz(3 * (size_t)n), b((size_t)nrhs * n), x((size_t)nrhs * n);
    if (!get(f, rp.data(), rp.size()) || !get(f, col.data(), col.size()) || !get(f, val.data(), val.size()) ||
        !get(f, xyz.data(), xyz.size()) || !get(f, b.data(), b.size()))
        return 2;
    double worst = 0.0;
    for (int budget : {-1, 0, 128, 1024})
        for (int emu = 0; emu < 2; ++emu)
            for (int pairs = 0; pairs < 2; ++pairs) {
                int64_t pe = 0;
                int lv = 0;
                if (md_host_factor_solve(n, rp.data(), col.data(), val.data(), xyz.data(), nrhs, b.data(), x.data(), &pe,
                                         &lv, budget, emu, pairs) != MD_OK)
                    return 1;
                for (int r = 0; r < nrhs; ++r) {
                    double num = 0.0, den = 0.0;
                    for (int i = 0; i < n; ++i) {
                        double y = 0.0;
                        for (int e = rp[i]; e < rp[i + 1]; ++e) y += val[e] * x[(size_t)r * n + col[e]];
                        num = std::fmax(num, std::fabs(y - b[(size_t)r * n + i]));
                        den = std::fmax(den, std::fabs(b[(size_t)r * n + i]));
                    }
                    worst = std::fmax(worst, num / (den > 0 ? den : 1.0));
                }
            }
    std::printf("system residual %.3e\n", worst);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: test_host_asan mesh|system FILE\n"); return 2; }
    const std::string mode = argv[1];
    if (mode == "mesh") return run_mesh(argv[2]);
    if (mode == "system") return run_system(argv[2]);
    return 2;
}
