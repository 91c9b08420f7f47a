// HIP kernels of the ARAP path for CDNA4 (gfx950): cotan assembly, local step, rotated-edge
// right-hand side with FP64 residual, level-scheduled block-LDL^T sweeps, ARAP energy.
//
// Every kernel works on a LaunchSet (one or many meshes); the mesh of a block is found by a
// binary search of the block index in the launch's prefix table, so a single mesh, a static
// batch and a served mix of meshes all run through the same launches.
#include "md_kernels.h"
#include "md_analysis.h"

namespace md {

namespace {

__device__ __forceinline__ int find_mesh(const LaunchSet &L, int b) {
    int lo = 0, hi = L.nactive - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (L.prefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename T> __device__ __forceinline__ T tsqrt(T x);
template <> __device__ __forceinline__ float tsqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double tsqrt<double>(double x) { return sqrt(x); }

// std::max<Scalar>(a, b) = (a < b) ? b : a; arap.h:199-218 pass the clamp first, so NaN
// yields the clamp.
__device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }

// ---- cotan weights (arap.h:182-239) -----------------------------------------------------

// One thread per face: the three clamped half-cotangents of arap.h:199-218, 225-227.
__global__ __launch_bounds__(kBlock) void k_cotan_face(LaunchSet L) {
#pragma clang fp contract(off)
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int f = (blockIdx.x - L.prefix[k]) * kBlock + threadIdx.x;
    if (f >= M.F) return;
    int i0 = M.tri[3 * f], i1 = M.tri[3 * f + 1], i2 = M.tri[3 * f + 2];
    d4 v0 = M.p[i0], v1 = M.p[i1], v2 = M.p[i2];
    double dx, dy, dz;
    dx = v1.x - v0.x; dy = v1.y - v0.y; dz = v1.z - v0.z;
    double l0 = ref_max(1e-8, dx * dx + dy * dy + dz * dz);
    dx = v2.x - v1.x; dy = v2.y - v1.y; dz = v2.z - v1.z;
    double l1 = ref_max(1e-8, dx * dx + dy * dy + dz * dz);
    dx = v0.x - v2.x; dy = v0.y - v2.y; dz = v0.z - v2.z;
    double l2 = ref_max(1e-8, dx * dx + dy * dy + dz * dz);
    l0 = sqrt(l0); l1 = sqrt(l1); l2 = sqrt(l2);
    const double semip = 0.5 * (l0 + l1 + l2);
    const double area = ref_max(1e-8, sqrt(semip * (semip - l0) * (semip - l1) * (semip - l2)));
    const double denom = 1.0 / (4.0 * area);
    double cot0 = (-l0 * l0 + l1 * l1 + l2 * l2) * denom;
    double cot1 = (l0 * l0 - l1 * l1 + l2 * l2) * denom;
    double cot2 = (l0 * l0 + l1 * l1 - l2 * l2) * denom;
    M.fe_val[3 * f] = ref_max(1e-10, cot0) * 0.5;
    M.fe_val[3 * f + 1] = ref_max(1e-10, cot1) * 0.5;
    M.fe_val[3 * f + 2] = ref_max(1e-10, cot2) * 0.5;
}

// One lane per vertex, one wavefront per 64-vertex slice: each slot sums its face-edge
// contributions in face order (setFromTriplets' duplicate sum, arap.h:224,238) and stores
// w_ij and the weighted rest edge a_ij = w_ij (p_i - p_j) in FP64 and FP32.
__global__ __launch_bounds__(kBlock) void k_cotan_slot(LaunchSet L) {
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices) return;
    int iv = slice * kSlice + lane;
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    d4 pi = M.p[iv < M.n ? iv : 0];
    for (int q = 0; q < wd; ++q) {
        long long g = (long long)(s0 + q) * kSlice + lane;
        double w = 0.0;
        for (int e = M.slot_fe_ptr[g]; e < M.slot_fe_ptr[g + 1]; ++e) w += M.fe_val[M.slot_fe[e]];
        int j = M.sell_col[g];
        d4 pj = M.p[j];
        double ax = w * (pi.x - pj.x), ay = w * (pi.y - pj.y), az = w * (pi.z - pj.z);
        if (iv >= M.n) { w = 0.0; ax = ay = az = 0.0; }
        M.w64[g] = w;
        M.a64[g] = ax; M.a64[M.slots + g] = ay; M.a64[2 * M.slots + g] = az;
        M.a32[g] = (float)ax; M.a32[M.slots + g] = (float)ay; M.a32[2 * M.slots + g] = (float)az;
    }
}

// ---- local step (arap.h:354-384) ----------------------------------------------------------

// One Jacobi rotation of the symmetric 4x4 matrix a on the pair (P, Q); v accumulates the
// eigenvectors (columns).
template <typename T, int P, int Q>
__device__ __forceinline__ void jrot(T (&a)[4][4], T (&v)[4][4]) {
    T apq = a[P][Q];
    T scale = fabs(a[P][P]) + fabs(a[Q][Q]);
    if (!(fabs(apq) > (sizeof(T) == 4 ? T(1e-9) : T(1e-18)) * scale) ) {
        if (fabs(apq) <= (sizeof(T) == 4 ? T(1e-30) : T(1e-300))) return;
        if (scale > T(0)) return;
    }
    T theta = (a[Q][Q] - a[P][P]) / (T(2) * apq);
    T t = (theta >= T(0) ? T(1) : T(-1)) / (fabs(theta) + tsqrt<T>(theta * theta + T(1)));
    T c = T(1) / tsqrt<T>(t * t + T(1));
    T s = t * c;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r == P || r == Q) continue;
        T arp = a[r][P], arq = a[r][Q];
        T np = c * arp - s * arq, nq = s * arp + c * arq;
        a[r][P] = np; a[P][r] = np;
        a[r][Q] = nq; a[Q][r] = nq;
    }
    a[P][P] = a[P][P] - t * apq;
    a[Q][Q] = a[Q][Q] + t * apq;
    a[P][Q] = T(0); a[Q][P] = T(0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        T vp = v[r][P], vq = v[r][Q];
        v[r][P] = c * vp - s * vq;
        v[r][Q] = s * vp + c * vq;
    }
}

// R = argmax_{R in SO(3)} tr(R M) for the covariance M = sum_j a_j e'_j^T, returned as the
// unit quaternion (x, y, z, w). This is the rotation V diag(1,1,det(V U^T)) U^T of the
// reference's JacobiSVD step (arap.h:376-382), computed as the dominant eigenvector of
// Horn's symmetric 4x4 matrix (J. Opt. Soc. Am. A 4(4), 1987) by cyclic Jacobi sweeps.
template <typename T>
__device__ __forceinline__ void best_rotation(const T (&m)[9], T &qx, T &qy, T &qz, T &qw) {
    const T Sxx = m[0], Sxy = m[1], Sxz = m[2], Syx = m[3], Syy = m[4], Syz = m[5], Szx = m[6],
            Szy = m[7], Szz = m[8];
    T a[4][4] = {{Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx},
                 {Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz},
                 {Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy},
                 {Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz}};
    T v[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    constexpr int kSweeps = sizeof(T) == 4 ? 6 : 10;
    for (int sw = 0; sw < kSweeps; ++sw) {
        T off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[0][3]) + fabs(a[1][2]) + fabs(a[1][3]) + fabs(a[2][3]);
        T dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]) + fabs(a[3][3]);
        if (!(off > (sizeof(T) == 4 ? T(1e-8) : T(1e-17)) * dia)) break;
        jrot<T, 0, 1>(a, v); jrot<T, 0, 2>(a, v); jrot<T, 0, 3>(a, v);
        jrot<T, 1, 2>(a, v); jrot<T, 1, 3>(a, v); jrot<T, 2, 3>(a, v);
    }
    T best = a[0][0];
    T q0 = v[0][0], q1 = v[1][0], q2 = v[2][0], q3 = v[3][0];
    if (a[1][1] > best) { best = a[1][1]; q0 = v[0][1]; q1 = v[1][1]; q2 = v[2][1]; q3 = v[3][1]; }
    if (a[2][2] > best) { best = a[2][2]; q0 = v[0][2]; q1 = v[1][2]; q2 = v[2][2]; q3 = v[3][2]; }
    if (a[3][3] > best) { best = a[3][3]; q0 = v[0][3]; q1 = v[1][3]; q2 = v[2][3]; q3 = v[3][3]; }
    T nrm = T(1) / tsqrt<T>(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    // canonical sign: scalar part >= 0
    if (q0 < T(0)) nrm = -nrm;
    qw = q0 * nrm; qx = q1 * nrm; qy = q2 * nrm; qz = q3 * nrm;
}

template <typename T>
__device__ __forceinline__ void quat_to_mat(T qx, T qy, T qz, T qw, T (&R)[9]) {
    T s = T(2) / (qx * qx + qy * qy + qz * qz + qw * qw);
    R[0] = T(1) - s * (qy * qy + qz * qz); R[1] = s * (qx * qy - qw * qz); R[2] = s * (qx * qz + qw * qy);
    R[3] = s * (qx * qy + qw * qz); R[4] = T(1) - s * (qx * qx + qz * qz); R[5] = s * (qy * qz - qw * qx);
    R[6] = s * (qx * qz - qw * qy); R[7] = s * (qy * qz + qw * qx); R[8] = T(1) - s * (qx * qx + qy * qy);
}

template <typename T> __device__ __forceinline__ const T *aplanes(const MeshDev &M);
template <> __device__ __forceinline__ const float *aplanes<float>(const MeshDev &M) { return M.a32; }
template <> __device__ __forceinline__ const double *aplanes<double>(const MeshDev &M) { return M.a64; }

// One lane per vertex (all n, constrained ones too, as arap.h:358 loops over every vertex):
// covariance sum_j w_ij (p_i - p_j)(p'_i - p'_j)^T (arap.h:366-374) from the slice's slots,
// then the rotation. The deformed edge p'_i - p'_j is formed in FP64 and rounded to T.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_local(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.n) return;
    const T *A = aplanes<T>(M);
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    d4 pi = M.pp[iv];
    T c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int q = 0; q < wd; ++q) {
        long long g = (long long)(s0 + q) * kSlice + lane;
        int j = M.sell_col[g];
        T ax = A[g], ay = A[M.slots + g], az = A[2 * M.slots + g];
        d4 pj = M.pp[j];
        T ex = (T)(pi.x - pj.x), ey = (T)(pi.y - pj.y), ez = (T)(pi.z - pj.z);
        c[0] += ax * ex; c[1] += ax * ey; c[2] += ax * ez;
        c[3] += ay * ex; c[4] += ay * ey; c[5] += ay * ez;
        c[6] += az * ex; c[7] += az * ey; c[8] += az * ez;
    }
    T qx, qy, qz, qw;
    best_rotation<T>(c, qx, qy, qz, qw);
    V out;
    out.x = qx; out.y = qy; out.z = qz; out.w = qw;
    reinterpret_cast<V *>(M.quat)[iv] = out;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_identity_quat(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int iv = (blockIdx.x - L.prefix[k]) * kBlock + threadIdx.x;
    if (iv >= M.n) return;
    V out;
    out.x = 0; out.y = 0; out.z = 0; out.w = 1;
    reinterpret_cast<V *>(M.quat)[iv] = out;
}

// ---- right-hand side and residual (arap.h:393-414) --------------------------------------

// One lane per free vertex i:
//   r_i = sum_j [ 1/2 (R_i + R_j) a_ij  -  w_ij (p'_i - p'_j) ]
// over all neighbours j. With p'_j equal to the targets for constrained j this is
// b_i - (L p')_i for the reference's b = bFixed + sum_j w_ij/2 (R_i + R_j)(p_i - p_j)
// (arap.h:395-414) and L of arap.h:325-332, so the solve returns the correction to p'.
// The rotation term is in T, the weight term in FP64.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices_free) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.nf) return;
    const T *A = aplanes<T>(M);
    const V *Q = reinterpret_cast<const V *>(M.quat);
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    d4 pi = M.pp[iv];
    V qi = Q[iv];
    T sax = 0, say = 0, saz = 0;   // sum_j a_ij
    T rx = 0, ry = 0, rz = 0;      // sum_j R_j a_ij
    double lx = 0, ly = 0, lz = 0; // sum_j w_ij (p'_i - p'_j)
    for (int q = 0; q < wd; ++q) {
        long long g = (long long)(s0 + q) * kSlice + lane;
        int j = M.sell_col[g];
        T ax = A[g], ay = A[M.slots + g], az = A[2 * M.slots + g];
        double w = M.w64[g];
        V qj = Q[j];
        d4 pj = M.pp[j];
        T Rj[9];
        quat_to_mat<T>(qj.x, qj.y, qj.z, qj.w, Rj);
        sax += ax; say += ay; saz += az;
        rx += Rj[0] * ax + Rj[1] * ay + Rj[2] * az;
        ry += Rj[3] * ax + Rj[4] * ay + Rj[5] * az;
        rz += Rj[6] * ax + Rj[7] * ay + Rj[8] * az;
        lx += w * (pi.x - pj.x); ly += w * (pi.y - pj.y); lz += w * (pi.z - pj.z);
    }
    T Ri[9];
    quat_to_mat<T>(qi.x, qi.y, qi.z, qi.w, Ri);
    T bx = T(0.5) * (Ri[0] * sax + Ri[1] * say + Ri[2] * saz + rx);
    T by = T(0.5) * (Ri[3] * sax + Ri[4] * say + Ri[5] * saz + ry);
    T bz = T(0.5) * (Ri[6] * sax + Ri[7] * say + Ri[8] * saz + rz);
    V out;
    out.x = (T)((double)bx - lx);
    out.y = (T)((double)by - ly);
    out.z = (T)((double)bz - lz);
    out.w = 0;
    reinterpret_cast<V *>(M.r)[iv] = out;
}

// ---- block LDL^T sweeps (arap.h:416-429) --------------------------------------------------

template <typename T> __device__ __forceinline__ const T *panels(const MeshDev &M) {
    return reinterpret_cast<const T *>(M.panel);
}

// Forward sweep, one work item = (supernode s, 64 off-diagonal rows from r0):
//   v = r[c0 .. c0+w) + children's update vectors (extend-add through rel, child order),
//   z_s = v_d (first item only),  u_s[r0 .. r0+64) = v_o - G_s v_d.
// The 4 wavefronts split the w columns of G_s; partial sums meet in LDS.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fwd(LaunchSet L, int level) {
    using V = typename V4T<T>::type;
    __shared__ T vd[3][kMaxSnWidth];
    __shared__ T vo[3][kFwdRows];
    __shared__ T part[4][3][kFwdRows];
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int item = M.fwd_ptr[level] + (blockIdx.x - L.prefix[k]);
    int s = M.fwd_s[item], r0 = M.fwd_r0[item];
    int c0 = M.sn_c0[s], w = M.sn_w[s], no = M.sn_noff[s];
    int m = w + no;
    int nrows = min(kFwdRows, no - r0);
    const V *R = reinterpret_cast<const V *>(M.r);
    const V *U = reinterpret_cast<const V *>(M.u);
    int tid = threadIdx.x;
    for (int p = tid; p < w; p += kBlock) {
        V b = R[c0 + p];
        vd[0][p] = b.x; vd[1][p] = b.y; vd[2][p] = b.z;
    }
    if (tid < kFwdRows) { vo[0][tid] = 0; vo[1][tid] = 0; vo[2][tid] = 0; }
    __syncthreads();
    for (int ci = M.child_ptr[s]; ci < M.child_ptr[s + 1]; ++ci) {
        int c = M.child[ci];
        int cn = M.sn_noff[c], cr = M.sn_roff[c];
        for (int e = tid; e < cn; e += kBlock) {
            int pos = M.rel[cr + e];
            V uv = U[cr + e];
            if (pos < w) {
                vd[0][pos] += uv.x; vd[1][pos] += uv.y; vd[2][pos] += uv.z;
            } else {
                int t = pos - w - r0;
                if (t >= 0 && t < nrows) { vo[0][t] += uv.x; vo[1][t] += uv.y; vo[2][t] += uv.z; }
            }
        }
        __syncthreads();
    }
    if (r0 == 0) {
        V *Z = reinterpret_cast<V *>(M.z);
        for (int p = tid; p < w; p += kBlock) {
            V zz;
            zz.x = vd[0][p]; zz.y = vd[1][p]; zz.z = vd[2][p]; zz.w = 0;
            Z[c0 + p] = zz;
        }
    }
    if (nrows <= 0) return;
    int wave = tid >> 6, lane = tid & 63;
    const T *P = panels<T>(M) + M.sn_poff[s] + w + r0;
    int qb = (w * wave) / 4, qe = (w * (wave + 1)) / 4;
    T ax = 0, ay = 0, az = 0;
    if (lane < nrows) {
        for (int q = qb; q < qe; ++q) {
            T g = P[(long long)q * m + lane];
            ax += g * vd[0][q]; ay += g * vd[1][q]; az += g * vd[2][q];
        }
    }
    part[wave][0][lane] = ax; part[wave][1][lane] = ay; part[wave][2][lane] = az;
    __syncthreads();
    if (wave == 0 && lane < nrows) {
        V out;
        out.x = vo[0][lane] - (part[0][0][lane] + part[1][0][lane] + part[2][0][lane] + part[3][0][lane]);
        out.y = vo[1][lane] - (part[0][1][lane] + part[1][1][lane] + part[2][1][lane] + part[3][1][lane]);
        out.z = vo[2][lane] - (part[0][2][lane] + part[1][2][lane] + part[2][2][lane] + part[3][2][lane]);
        out.w = 0;
        reinterpret_cast<V *>(M.u)[M.sn_roff[s] + r0 + lane] = out;
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Backward sweep, one work item = (supernode s, 16 columns from q0):
//   x_s[q] = sum_p Dinv_s[p, q] z_s[p] - sum_r G_s[r, q] x[offrow_r]
// (the panel's column q dotted with [z_s; -x_off]); the correction is added to p'.
// Each wavefront owns 4 columns; lanes stride the rows; rows are staged through LDS.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bwd(LaunchSet L, int level) {
    using V = typename V4T<T>::type;
    constexpr int kStage = 1024;
    __shared__ T tv[3][kStage];
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int item = M.bwd_ptr[level] + (blockIdx.x - L.prefix[k]);
    int s = M.bwd_s[item], q0 = M.bwd_q0[item];
    int c0 = M.sn_c0[s], w = M.sn_w[s], no = M.sn_noff[s], roff = M.sn_roff[s];
    int m = w + no;
    int ncols = min(kBwdCols, w - q0);
    const V *Z = reinterpret_cast<const V *>(M.z);
    V *X = reinterpret_cast<V *>(M.x);
    const T *P = panels<T>(M) + M.sn_poff[s];
    int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    T acc[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) { acc[i][0] = 0; acc[i][1] = 0; acc[i][2] = 0; }
    for (int pb = 0; pb < m; pb += kStage) {
        int cnt = min(kStage, m - pb);
        for (int p = tid; p < cnt; p += kBlock) {
            int gp = pb + p;
            V t;
            if (gp < w) {
                t = Z[c0 + gp];
            } else {
                t = X[M.offrows[roff + gp - w]];
                t.x = -t.x; t.y = -t.y; t.z = -t.z;
            }
            tv[0][p] = t.x; tv[1][p] = t.y; tv[2][p] = t.z;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int qc = wave + 4 * i;
            if (qc < ncols) {
                const T *col = P + (long long)(q0 + qc) * m + pb;
                for (int p = lane; p < cnt; p += 64) {
                    T g = col[p];
                    acc[i][0] += g * tv[0][p]; acc[i][1] += g * tv[1][p]; acc[i][2] += g * tv[2][p];
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int qc = wave + 4 * i;
        T sx = wave_sum<T>(acc[i][0]), sy = wave_sum<T>(acc[i][1]), sz = wave_sum<T>(acc[i][2]);
        if (qc < ncols && lane == 0) {
            int col = c0 + q0 + qc;
            V out;
            out.x = sx; out.y = sy; out.z = sz; out.w = 0;
            X[col] = out;
            d4 pv = M.pp[col];
            pv.x += (double)sx; pv.y += (double)sy; pv.z += (double)sz;
            M.pp[col] = pv;
        }
    }
}

// ---- ARAP energy -------------------------------------------------------------------------

// E = sum_i sum_j w_ij |(p'_i - p'_j) - R_i (p_i - p_j)|^2 over all vertices, in FP64; one
// partial sum per block of 256 vertices, reduced in a fixed order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy(LaunchSet L) {
    using V = typename V4T<T>::type;
    __shared__ double red[kBlock];
    int k = find_mesh(L, blockIdx.x);
    const MeshDev &M = L.meshes[k];
    int lb = blockIdx.x - L.prefix[k];
    int slice = lb * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    int iv = slice * kSlice + lane;
    double e = 0.0;
    if (slice < M.nslices && iv < M.n) {
        int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
        V qi = reinterpret_cast<const V *>(M.quat)[iv];
        double R[9];
        quat_to_mat<double>((double)qi.x, (double)qi.y, (double)qi.z, (double)qi.w, R);
        d4 pi = M.p[iv], ppi = M.pp[iv];
        for (int q = 0; q < wd; ++q) {
            long long g = (long long)(s0 + q) * kSlice + lane;
            int j = M.sell_col[g];
            double w = M.w64[g];
            d4 pj = M.p[j], ppj = M.pp[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            double tx = (ppi.x - ppj.x) - (R[0] * dx + R[1] * dy + R[2] * dz);
            double ty = (ppi.y - ppj.y) - (R[3] * dx + R[4] * dy + R[5] * dz);
            double tz = (ppi.z - ppj.z) - (R[6] * dx + R[7] * dy + R[8] * dz);
            e += w * (tx * tx + ty * ty + tz * tz);
        }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) M.energy_part[lb] = red[0];
}

}  // namespace

void launch_cotan_face(hipStream_t s, LaunchSet L, int blocks) {
    if (blocks > 0) hipLaunchKernelGGL(k_cotan_face, dim3(blocks), dim3(kBlock), 0, s, L);
}
void launch_cotan_slot(hipStream_t s, LaunchSet L, int blocks) {
    if (blocks > 0) hipLaunchKernelGGL(k_cotan_slot, dim3(blocks), dim3(kBlock), 0, s, L);
}
void launch_local(hipStream_t s, LaunchSet L, int blocks, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_local<double>, dim3(blocks), dim3(kBlock), 0, s, L);
    else hipLaunchKernelGGL(k_local<float>, dim3(blocks), dim3(kBlock), 0, s, L);
}
void launch_identity_quat(hipStream_t s, LaunchSet L, int blocks, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_identity_quat<double>, dim3(blocks), dim3(kBlock), 0, s, L);
    else hipLaunchKernelGGL(k_identity_quat<float>, dim3(blocks), dim3(kBlock), 0, s, L);
}
void launch_residual(hipStream_t s, LaunchSet L, int blocks, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_residual<double>, dim3(blocks), dim3(kBlock), 0, s, L);
    else hipLaunchKernelGGL(k_residual<float>, dim3(blocks), dim3(kBlock), 0, s, L);
}
void launch_fwd(hipStream_t s, LaunchSet L, int blocks, int level, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_fwd<double>, dim3(blocks), dim3(kBlock), 0, s, L, level);
    else hipLaunchKernelGGL(k_fwd<float>, dim3(blocks), dim3(kBlock), 0, s, L, level);
}
void launch_bwd(hipStream_t s, LaunchSet L, int blocks, int level, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_bwd<double>, dim3(blocks), dim3(kBlock), 0, s, L, level);
    else hipLaunchKernelGGL(k_bwd<float>, dim3(blocks), dim3(kBlock), 0, s, L, level);
}
void launch_energy(hipStream_t s, LaunchSet L, int blocks, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_energy<double>, dim3(blocks), dim3(kBlock), 0, s, L);
    else hipLaunchKernelGGL(k_energy<float>, dim3(blocks), dim3(kBlock), 0, s, L);
}

}  // namespace md
