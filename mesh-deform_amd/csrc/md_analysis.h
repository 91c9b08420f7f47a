// Host-side analysis of one ARAP mesh: vertex renumbering, one-ring slices, fill-reducing
// ordering and the supernodal block-LDL^T factor consumed by the HIP solve kernels.
//
// Replaces, for the path of inc/deform/arap.h (cheind/mesh-deform):
//   initializeFreeVariableMapping (arap.h:261-272) -> free vertices first in solve order,
//   setupLinearSystem (arap.h:292-340)             -> supernodal symbolic + FP64 numeric
//                                                      factorization (Eigen SimplicialLDLT
//                                                      with AMD ordering in the reference).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace md {

// Maximum columns of one supernode (wider separators are split into chains).
constexpr int kMaxSnWidth = 1024;
// Maximum columns of a supernode merged from a level pair (a supernode with its children).
constexpr int kMaxGroupWidth = 4096;
// Subtree budget whose sweep levels the level pairing follows, whatever budget the layout then
// uses: the factor depends on the amalgamation and the pairing only, so a re-layout with any
// budget keeps it (and the small budget pairs the low levels as well).
constexpr int kPairBudget = 128;
// Edge of the square panel tile one wavefront streams in the triangular sweeps.
constexpr int kTile = 64;
// Vertices per one-ring slice (one wavefront).
constexpr int kSlice = 64;
// Default LDS budget (vector entries) of one bottom subtree swept by a single wavefront.
constexpr int kSubBudgetDefault = 512;
// Default relaxed-amalgamation limit: a child merges into its parent while the explicit zeros
// stay below this share of the merged panel (half of it for merges wider than 48 columns).
constexpr double kAmalgRatioDefault = 0.3;
// Largest budget: 1024 FP64 vector entries (32 bytes each) are 32 KiB of LDS; with at most
// kSubMaxNodes 64-byte supernode records a subtree stays within 36 KiB.
constexpr int kSubBudgetMax = 1024;
constexpr int kSubMaxNodes = 64;
// Largest supernode a bottom subtree may hold: columns, and rows of the front.
constexpr int kSubMaxWidth = 64;
constexpr int kSubMaxRows = 128;

// One-ring topology in the order the kernels use ("internal" vertex ids):
//   internal ids [0, nf) are the free vertices in solve (elimination) order,
//   internal ids [nf, n) are the constrained vertices in ascending original id.
// Each slice of 64 internal vertices stores its neighbours column by column
// (slot k of lane l at (slice_ptr[s] + k) * 64 + l), padded with the vertex itself.
struct Topology {
    int n = 0, F = 0, nf = 0;
    std::vector<int> internal_of;   // original id -> internal id
    std::vector<int> original_of;   // internal id -> original id
    int nslices = 0;
    std::vector<int> slice_ptr;     // nslices + 1, in 64-wide slot columns
    std::vector<int> sell_col;      // slots * 64, internal neighbour id (self for padding)
    std::vector<int> sell_orig_rank;// not stored on device: for the weight build
    // cotan build: each face edge (3 per face) writes its half-cotangent to fe_val[3f+k];
    // each slot sums the face edges listed in slot_fe[slot_fe_ptr[slot] ..] in face order.
    std::vector<int> tri_internal;  // 3F internal ids
    std::vector<int> slot_fe_ptr;   // total_slots + 1
    std::vector<int> slot_fe;       // face-edge ids (3f + k)
    int64_t total_slots() const { return (int64_t)slice_ptr[nslices] * kSlice; }
};

// Supernodal block LDL^T: A = Lb * D * Lb^T with Lb unit block-lower-triangular.
// For supernode s with columns [c0, c0 + w) and off-diagonal rows offrows[roff .. roff+noff)
// the panel (m = w + noff rows, w columns, column-major) holds
//   rows [0, w):  Dinv_s = D_s^{-1} (dense symmetric),
//   rows [w, m):  G_s = L_os L_ss^{-1} (the block of Lb below D_s).
// Forward:  v = b[c0..] + children updates; z_s = v_d; u_s = v_o - G_s v_d.
// Backward: x_s = Dinv_s z_s - G_s^T x[offrows].
//
// On the device both sweeps are cut into 64 x 64 tiles, one wavefront each:
//   forward tile (s, r0, q0): rows [r0, r0+64) of G_s times columns [q0, q0+64) of v_d,
//     read from a column-major copy of G_s; it writes partial kc = q0/64 of u_s, so u_s is
//     stored as kc_s = ceil(w/64) partial vectors that the parent sums in order;
//   backward tile (s, q0, p0): rows [p0, p0+64) of [Dinv_s; G_s]^T [z_s; -x_off], read
//     from a row-major copy; it writes partial kb = p0/64 of x_s, so x_s is stored as
//     kr_s = ceil(m/64) partials that every reader (descendants, the final update) sums.
//
// Bottom subtrees: a subtree whose supernodes are all small (w <= 64, m <= 128) and whose
// fronts and columns fit an LDS budget is swept by one wavefront, height by height, with
// every front in LDS (no level launches, no global round trips inside the subtree). Its
// supernodes keep one partial of x_s (kr = 1); only the subtree root stores u_s (kc = 1,
// the others kc = 0). Each supernode of a subtree has a 16-int record (SubNode layout:
// c0, w, noff, forward front base, G offset, [Dinv; G] offset, contribution base, off-row
// source base, u output or -1, x output, LDS column base of x, position / off-row / column
// prefix within its height, backward front base, 0). Records of a subtree are contiguous
// and ordered by height; the wavefront copies them to LDS and maps lane items to supernodes
// by binary search over the prefixes. LDS holds, forward, the fronts of all the subtree's
// supernodes; backward, the solution columns of the subtree and the fronts of one height.
// Per item only this remains in HBM:
//   forward positions: the LDS indices of the first two children's update entries landing
//     there (-1: none); a second index <= -2 points (-2 - i) at sfc[i ..], a -1-terminated
//     list of the rest;
//   backward off-row positions: (LDS index of x, -1) when the row lies in the subtree, else
//     (first partial of x, count << 16 | stride) of the row's owner.
// Level schedule: supernodes outside subtrees get level = 1 + the largest level of their
// children outside subtrees (0 if none); subtree sweeps run before the forward levels and
// after the backward levels.
struct Factor {
    int n = 0;                       // dimension (free vertices)
    int ns = 0;                      // supernodes
    std::vector<int> c0, w, noff;    // per supernode
    std::vector<int> parent;         // -1 for roots
    std::vector<int> height;         // 0 for leaves
    int nlevels = 0;
    std::vector<int> roff;           // ns + 1: offsets into offrows / rel
    std::vector<int> offrows;        // ascending, solve numbering
    std::vector<int> rel;            // per off row of s: position in parent's rows [0, m_p)
    std::vector<int> child_ptr, child;
    std::vector<int64_t> poff;       // ns + 1: FP64 panel offsets (m x w, column-major)
    std::vector<double> panel;       // FP64 panels (host factorization)
    // device numeric factorization (multifrontal, one height at a time): per height the
    // supernodes of that height; per supernode the offsets of its workspace blocks: the
    // off-diagonal front block (noff x w), the inverse of the pivot block's Cholesky factor
    // (w x w) and the update matrix for the parent (noff x noff), all column-major FP64
    std::vector<int> hptr, hlist;    // height h: hlist[hptr[h] .. hptr[h+1])
    std::vector<int64_t> woff_s, woff_t, woff_u;  // ns + 1 each
    std::vector<int> col_sn;         // supernode of each column
    // device layout
    std::vector<int> kc, kr;         // partial counts of u_s and x_s
    std::vector<int64_t> pc_off;     // ns + 1: column-major G_s (noff x w)
    std::vector<int64_t> pr_off;     // ns + 1: row-major [Dinv_s; G_s] (m x w)
    // update slots of the forward sweep: a level supernode's front has uk slots per position
    // (slot k of position p at ubase + k * m + p, k-major); the children's partial updates for
    // p fill consecutive slots in child order (kc each), the others stay zero, so a front sums
    // its slots at addresses known from its tile alone
    std::vector<int64_t> ubase;      // ns + 1: first slot of each supernode's front
    std::vector<int> uk;             // per supernode: slots per front position
    std::vector<int64_t> udst;       // per off-row entry of a level supernode or subtree root:
                                     //   its slot in the parent's front (partial cb: + cb * m_parent)
    std::vector<int64_t> xoff;       // ns + 1: kr * w partial solution entries
    std::vector<int> rowinfo;        // 4 per row: first partial, stride, count, 0
    std::vector<int> xloc;           // 4 per off row entry: rowinfo of that row
    // tiles grouped by height level: level l occupies [ft_ptr[l], ft_ptr[l+1]) (4 ints each)
    std::vector<int> ft_ptr, ft;     // forward (s, r0, q0, 0)
    std::vector<int> bt_ptr, bt;     // backward (s, q0, p0, 0)
    std::vector<int> level;          // device level of supernodes outside subtrees, -1 inside
    // bottom subtrees
    int nsub = 0;                    // subtrees
    int sub_lds_fwd = 0, sub_lds_bwd = 0;  // LDS vector entries of the largest, per sweep
    int sub_nodes = 0;               // supernodes of the largest subtree
    std::vector<int> sub;            // 4 per subtree: (first record, end record, height table offset, heights)
    std::vector<int> sub_h;          // per (subtree, height) + a terminal per subtree: first record
    std::vector<int> snode;          // 16 per subtree supernode (SubNode), by subtree then height
    std::vector<int> scp;            // 2 per forward position: first two child entries (see above)
    std::vector<int> sfc;            // -1-terminated overflow lists of child entries
    std::vector<int> sbo;            // 2 per off-row position of a subtree supernode (backward source)
    std::vector<int> srow;           // 2 per subtree: first row, rows
    std::vector<int> sfi;            // per row (n): forward front slot of a subtree row, else -1
    int64_t nnz_exact = 0;           // strictly-lower nonzeros of the exact factor (no amalgamation zeros)
    int64_t factor_nnz() const { return poff.empty() ? 0 : poff.back(); }
};

struct AnalysisStats {
    int n, nf, nslices, ns, nlevels;
    int64_t slots, panel_entries, strict_nnz_equiv, nnz_exact, fwd_items, bwd_items;
    int max_width, max_rows;
};

// Builds the topology for faces `tri` (3F original ids) of an n-vertex mesh with the
// constrained flags `is_con` (n entries) and the solve order `perm` of the free vertices
// (perm[k] = original id of the k-th free vertex in elimination order).
void build_topology(int n, int F, const int* tri, const std::vector<int>& perm,
                    const std::vector<char>& is_con, Topology& topo);

// Fill-reducing order of the free vertices: geometric nested dissection on the free-free
// graph, followed by an elimination-tree postorder. Returns original vertex ids.
// With pair_group, the order also pairs sweep levels (pair_levels with kPairBudget and the
// given amalgamation) and pair_group receives the group of every position.
std::vector<int> order_free_vertices(int n, int F, const int* tri, const double* xyz,
                                     const std::vector<char>& is_con, std::vector<int>* pair_group = nullptr,
                                     double amalg_ratio = 0.0);

// The same ordering for a general symmetric pattern (full CSR, m rows) with coordinates.
std::vector<int> order_graph(int m, const int* rowptr, const int* col, const double* xyz);

// Symbolic supernodal analysis of the free-free graph in solve order.
// adj_ptr/adj (CSR over solve indices, no diagonal needed) describe A's pattern.
// sub_budget: LDS vector entries of one bottom subtree (0: no subtree sweeps); amalg_ratio:
// relaxed-amalgamation limit (more merges: fewer, wider supernodes and levels, more zeros).
// group (per column, -1: none): columns of one group that end up adjacent in the supernodal
// tree always merge (up to kMaxGroupWidth columns).
void symbolic(int n, const std::vector<int>& adj_ptr, const std::vector<int>& adj, Factor& f,
              int sub_budget = kSubBudgetDefault, double amalg_ratio = kAmalgRatioDefault,
              const std::vector<int>* group = nullptr);

// Level pairs: in the factor of this order (pattern ptr/adj over the original indices, laid
// out with the given subtree budget and amalgamation), each sweep level at an even distance
// below the top level joins the supernodes one level up that are its parents: the returned
// order places a group's member supernodes right before their parent (after the members'
// own subtrees, so the fill is unchanged) and group receives each position's group (-1:
// none), so that symbolic() merges every group into one supernode: about half the levels,
// for denser panels.
std::vector<int> pair_levels(const std::vector<int>& order, const std::vector<int>& ptr, const std::vector<int>& adj,
                             int sub_budget, double amalg_ratio, std::vector<int>& group);

// Device layout of a symbolic factor (partials, contribution lists, tiles, bottom subtrees)
// for the given subtree budget; symbolic() calls it, and it may be called again to re-lay
// the same structure out with another budget.
void device_layout(Factor& f, int sub_budget);

// FP64 multifrontal numeric factorization. A is given by its diagonal (diag[i]) and the
// strictly-off-diagonal entries adj/adj_val (CSR in solve order, symmetric pattern).
// Returns false (and sets err) when a pivot block is not positive definite.
bool numeric(const std::vector<int>& adj_ptr, const std::vector<int>& adj,
             const std::vector<double>& adj_val, const std::vector<double>& diag, Factor& f,
             std::string& err);

// The device numeric factorization emulated on the host (same steps and index math, every
// index checked); same inputs and result as numeric().
bool numeric_device_emulation(const std::vector<int>& adj_ptr, const std::vector<int>& adj,
                              const std::vector<double>& adj_val, const std::vector<double>& diag, Factor& f,
                              std::string& err);

// Host reference of the GPU solve on the same panels (used by host tests).
void solve_host(const Factor& f, const double* b, double* x, int nrhs);

// Tile-by-tile emulation of the device sweeps on 3 right-hand sides (b and x are 3 x n,
// coordinate-major), with every index checked against its buffer.
bool solve_tiles_host(const Factor& f, const double* b, double* x, std::string& err);

AnalysisStats stats(const Topology& t, const Factor& f);

}  // namespace md
