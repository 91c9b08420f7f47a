// Device-side data layout and kernel launchers of the ARAP path (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace md {

struct alignas(16) f4 { float x, y, z, w; };
struct alignas(32) d4 { double x, y, z, w; };

template <typename T> struct V4T;
template <> struct V4T<float> { using type = f4; };
template <> struct V4T<double> { using type = d4; };

// Per-supernode record of the block LDL^T sweeps (see Factor in md_analysis.h).
// Pointers of the device records below are in the global address space in the device pass of
// the kernels' file (md_kernels.hip defines MD_KERNELS_DEVICE), so the kernels' loads and stores
// are global (not flat) instructions: they wait on the vector memory counter alone, in issue
// order, instead of on the LDS counter as well. Host code sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__) && defined(MD_KERNELS_DEVICE)
#define MD_G __attribute__((address_space(1)))
#else
#define MD_G
#endif

struct SnDev {
    int c0, w, noff, roff;     // columns [c0, c0+w), off rows offrows[roff .. roff+noff)
    int kc, kr, uk, dstride;   // partial counts of u_s / x_s, update slots per front position,
                               //   the parent's front size (stride of u_s's partials in its slots)
    long long pc_off, pr_off;  // column-major G_s, row-major [Dinv_s; G_s]
    long long ubase, xoff;     // first update slot of the front / partial solution entries
};

// One level tile of a sweep: the supernode's record and the tile origin, so a tile is one load.
struct TileDev {
    SnDev s;
    int a, b, pad0, pad1;      // forward (r0, q0); backward (q0, p0)
};

// Device view of one mesh. All pointers are device pointers into the mesh's arena.
// Internal vertex ids: [0, nf) free vertices in solve order, [nf, n) constrained ones.
struct MeshDev {
    int n, nf, nslices, ns;
    int nslices_free;           // slices covering internal ids [0, nf)
    int F;
    long long slots;            // 64 * slice_ptr[nslices]
    const MD_G int *slice_ptr;  // nslices + 1 (64-wide slot columns)
    const MD_G int *sell_col;   // slots: neighbour internal id (self = padding)
    MD_G float *a32;            // 3 planes of slots: w_ij (p_i - p_j), FP32
    MD_G double *a64;           // 3 planes of slots, FP64
    MD_G double *w64;           // slots: cotan weight w_ij
    const MD_G int *tri;        // 3F internal ids
    MD_G double *fe_val;        // 3F half-cotangents (face edge 3f+k)
    const MD_G int *slot_fe_ptr; // slots + 1
    const MD_G int *slot_fe;    // face-edge ids feeding each slot, face order
    MD_G d4 *p;                 // rest positions (internal order)
    MD_G d4 *pp;                // current positions p'
    MD_G void *quat;            // rotations as quaternions (x, y, z, w=scalar): f4 or d4
    MD_G void *r, *z;           // residual and forward result (V4 of the solve type), nf each
    MD_G void *upart, *xpart;   // update slots of the level fronts / partial solutions (V4)
    const MD_G int *child_ptr, *child;
    const MD_G long long *udst; // per off-row entry of a level supernode or subtree root: its
                                //   update slot in the parent's front (partial cb: + cb * dstride)
    const MD_G int4 *rowinfo;   // per row: (first partial, stride, count, 0)
    const MD_G int4 *xloc;      // per off row entry: rowinfo of that row
    const MD_G void *pc, *pr;   // panels (float or double), forward and backward layouts
    const MD_G TileDev *ftile, *btile; // level tiles, level by level (bases in the launch tables)
    // bottom subtrees, one wavefront each (layouts in md_analysis.h, Factor)
    const MD_G int4 *sub;       // (first record, end record, height table offset, heights)
    const MD_G int *sub_h;      // first record of each (subtree, height), + a terminal per subtree
    const MD_G int4 *snode;     // supernode records, four int4 each (SubNode)
    const MD_G int2 *scp;       // forward positions: first two child update entries (LDS), overflow
    const MD_G int *sfc;        // -1-terminated overflow lists of child update entries
    const MD_G int2 *sbo;       // backward off-row sources
    int nsub;                   // bottom subtrees (their rows' x is added to p' by k_sub_bwd)
    const MD_G int2 *srow;      // per subtree: first row, rows (their residual is formed in LDS)
    const MD_G int *sfi;        // per free row of a subtree: its forward front slot in LDS
    int nlrows;                 // rows of level-swept supernodes, increasing
    const MD_G int *lrows;      //   (their residual: extra blocks of k_sub_fwd; x summed from the
                                //   partials and added to p': extra blocks of k_sub_bwd)
    MD_G double *energy_part;   // one partial sum per 256 vertices
    // Jacobi-PCG alternative to the triangular sweeps (solve type vectors, nf entries)
    MD_G float *w32;            // slots: cotan weights in FP32
    MD_G void *diag;            // nf: L_ii = sum_j w_ij (solve type scalar)
    MD_G void *cg_x, *cg_p;     // correction, direction (V4)
    MD_G void *cg_v;            // 5 vectors (V4): r (second buffer; the first is r), w and s (two each)
    MD_G void *cg_part;         // 4 nblk + 5 (V4): partials of (r, u) and (w, u) for two parities,
                                //   then (gamma, alpha) for two parities and the first gamma
    int nblk_free;              // 256-vertex blocks covering [0, nf)
    double pcg_tol2;            // stop (freeze) when r.z <= tol^2 * (r.z)_0, per coordinate
};

// Device numeric factorization (multifrontal, FP64): per supernode record and the buffers.
struct FacDev {
    int c0, w, noff, roff;          // columns, off rows (offrows[roff ..], rel[roff ..])
    int cb, ce, m, pad;             // children child[cb .. ce), m = w + noff
    long long poff, soff, toff, uoff;  // FP64 panel (m x w), F_os copy (noff x w), L^{-1} (w x w),
                                       // update matrix (noff x noff); all column-major
    long long pc_off, pr_off;       // solve-precision panels of the sweeps
};

struct FacArgs {
    const MD_G FacDev *sn;
    const MD_G int *hlist;          // supernodes by height
    const MD_G int *offrows, *rel, *child;
    const MD_G int *slice_ptr, *sell_col;
    const MD_G double *w64;         // cotan weights over the one-ring slots
    int nf;
    MD_G double *P, *S, *T, *U, *diag;
    MD_G unsigned long long *amax;  // max |L_ii| as the bits of a non-negative double
    MD_G int *flag;                 // 1: a pivot block is not positive definite; 2: structure
    MD_G void *pc, *pr;             // the sweeps' panels (float or double), written by the D^{-1}
    int solve64;                    //   and G kernels in the solve precision (1: double)
    MD_G double *Wk;                // partial products of the split-K launches (fac_split_k_doubles)
};

// A set of meshes processed by one launch. Block b belongs to mesh k where
// prefix[k] <= b < prefix[k+1]; `meshes` and `prefix` are device arrays.
struct LaunchSet {
    const MD_G MeshDev *meshes;
    const MD_G int *prefix;
    int nactive;
    const MD_G int *bmesh;      // optional: mesh of every block (static batches), else a binary search
    const MD_G int *tbase;      // per mesh: first tile of the launch's level (level sweeps), else 0
    const MD_G TileDev *tiles;  // one-mesh level launches: the level's first tile (else null)
};

// Launchers (grid = total blocks of the launch, from the host-side prefix table).
void launch_cotan_face(hipStream_t s, LaunchSet L, int blocks);
void launch_cotan_slot(hipStream_t s, LaunchSet L, int blocks);
void launch_local(hipStream_t s, LaunchSet L, int blocks, bool fp64);
void launch_residual(hipStream_t s, LaunchSet L, int blocks, bool fp64);
void launch_fwd(hipStream_t s, LaunchSet L, int blocks, int level, bool fp64);
void launch_bwd(hipStream_t s, LaunchSet L, int blocks, int level, bool fp64);
// Bottom-subtree sweeps: one 64-lane block per subtree; LDS for `nodes` supernode records
// and `lds_entries` vector entries.
// fused: the launch also forms the residual (one-mesh plans; blocks past the subtrees take
// the level rows), else k_residual ran before it.
void launch_sub_fwd(hipStream_t s, LaunchSet L, int blocks, int nodes, int lds_entries, bool fused, bool fp64);
void launch_sub_bwd(hipStream_t s, LaunchSet L, int blocks, int nodes, int lds_entries, bool fp64);
void launch_energy(hipStream_t s, LaunchSet L, int blocks, bool fp64);
void launch_identity_quat(hipStream_t s, LaunchSet L, int blocks, bool fp64);
// A target of a served request: the device address of its vertex's p' entry and the position.
struct alignas(16) TargetEntry {
    unsigned long long addr;
    double x, y, z;
};
// *entry.addr = (x, y, z) for k entries (targets of several meshes in one launch).
void launch_scatter_entries(hipStream_t s, const TargetEntry *e, int k);
void launch_normals(hipStream_t s, LaunchSet Lf, int blocks_face, LaunchSet Ls, int blocks_slice, double *fn, d4 *out);
void launch_pcg_diag(hipStream_t s, LaunchSet L, int blocks, bool fp64);
// Jacobi-PCG in the Chronopoulos-Gear form: one launch per CG iteration (step i).
void launch_pcg_init(hipStream_t s, LaunchSet L, int blocks, bool fp64);
void launch_pcg_step(hipStream_t s, LaunchSet L, int blocks, int step, bool fp64);
void launch_pcg_apply(hipStream_t s, LaunchSet L, int blocks, bool fp64);
// Device numeric factorization: diagonal and its maximum, then per height (the `count`
// supernodes from hlist[h0]) the front assembly and the pivot block's Cholesky and inverse (one
// workgroup per supernode), D^{-1}, G and the update matrix (one workgroup per output tile, or
// several split over K); D^{-1} and G also go to the solve-precision panels.
void launch_fac_diag(hipStream_t s, const FacArgs &A);
// A product stage of a height: workgroups per supernode (one per 64 x 64 output tile, lower tiles
// for the symmetric ones; the height's most) and splits over K (1: none); a split stage needs
// fac_split_k_doubles(count * per * ks) doubles at A.Wk.
struct FacSplit {
    int per, ks;
};
// One height: assembly (per_asm workgroups per supernode, one per asm_cols front columns), blocked
// Cholesky and inverse of the pivot blocks (nb = column blocks of 64 in the widest one), D^{-1},
// G and the update matrices (g.per = 0 when no supernode of the height has off rows).
void launch_fac_height(hipStream_t s, const FacArgs &A, int h0, int count, int nb, int asm_cols, int per_asm,
                       const FacSplit &dinv, const FacSplit &g, const FacSplit &u);
// Split of a product over K: `tiles` output tiles of 64 x 64 and `chunks` chunks of K (32 values).
// Splits only launches of few tiles (tiles x 2 <= kSplitWorkgroups) with at least 8 chunks, into
// shares of at least 4 chunks, at most kSplitWorkgroups workgroups in all.
constexpr int kSplitWorkgroups = 1024;
inline int fac_split_k(long long tiles, int chunks) {
    if (tiles <= 0 || tiles * 2 > kSplitWorkgroups || chunks < 8) return 1;
    const long long ks = std::min<long long>(chunks / 4, kSplitWorkgroups / tiles);
    return ks > 1 ? (int)ks : 1;
}
constexpr long long fac_split_k_doubles(long long workgroups) { return workgroups * 16 * 256; }
// A height whose fronts all have m <= mm <= kFacSmallFront and w <= mw: every step of
// launch_fac_height for one front in one wavefront (front in mm^2 + mm mw doubles of LDS: under
// 64 KiB), without the FP64 panel workspace.
constexpr int kFacSmallFront = 62;
void launch_fac_small(hipStream_t s, const FacArgs &A, int h0, int count, int mm, int mw);
// Internal-order vectors (positions, normals) to packed xyz in original order (perm: original
// id per internal id), and packed xyz to internal order (p and/or pp; pp takes the targets tgt,
// internal order, for the constrained ids [nf, n)).
void launch_pack_xyz(hipStream_t s, const d4 *v, const int *perm, int n, double *dst);
void launch_unpack_xyz(hipStream_t s, const double *xyz, const int *perm, int n, int nf, const double *tgt, d4 *p,
                       d4 *pp);

// Per-GPU device workspace of the non-allocating entry points (a static device array of
// kScratchBytes): its address on the current device, from hipGetSymbolAddress. Call once per
// device with that device current and use only the returned address in host code.
constexpr size_t kScratchBytes = size_t(64) << 20;
hipError_t scratch_address(void **ptr);
// Plain copy of n16 16-byte elements (counter calibration, copy bandwidth).
void launch_copy_probe(hipStream_t s, const void *src, void *dst, long long n16);

// Threads per block and work per block of each kernel (used for the prefix tables).
constexpr int kBlock = 256;
constexpr int kSlicesPerBlock = kBlock / 64;  // one 64-vertex slice per wavefront

}  // namespace md
