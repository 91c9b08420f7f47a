// HIP kernels of the ARAP path for CDNA4 (gfx950): cotan assembly, local step, rotated-edge
// right-hand side with FP64 residual, level-scheduled block-LDL^T sweeps, ARAP energy.
//
// Every kernel works on a LaunchSet (one or many meshes); the mesh of a block is found by a
// binary search of the block index in the launch's prefix table, so a single mesh, a static
// batch and a served mix of meshes all run through the same launches.
#define MD_KERNELS_DEVICE 1
#include "md_kernels.h"
#include "md_analysis.h"
#include "md_arap.h"

#include <cmath>
#include <string>

namespace md {

namespace {

__device__ __forceinline__ int find_mesh(const LaunchSet &L, int b) {
    if (L.bmesh) return L.bmesh[b];
    int lo = 0, hi = L.nactive - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (L.prefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Reciprocal, square root and reciprocal square root of the local step's Jacobi rotations:
// single hardware instructions in FP32 (about 1 ulp; the rotation is renormalised at the
// end), correctly rounded in FP64.
// (host builds of the rotation code, for the host diagnostic md_host_best_rotation, use the
// correctly rounded operations)
template <typename T> __host__ __device__ __forceinline__ T jrcp(T x);
template <> __host__ __device__ __forceinline__ float jrcp<float>(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
template <> __host__ __device__ __forceinline__ double jrcp<double>(double x) { return 1.0 / x; }
template <typename T> __host__ __device__ __forceinline__ T jsqrt(T x);
template <> __host__ __device__ __forceinline__ float jsqrt<float>(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}
template <> __host__ __device__ __forceinline__ double jsqrt<double>(double x) { return sqrt(x); }
template <typename T> __host__ __device__ __forceinline__ T jrsqrt(T x);
template <> __host__ __device__ __forceinline__ float jrsqrt<float>(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}
template <> __host__ __device__ __forceinline__ double jrsqrt<double>(double x) { return 1.0 / sqrt(x); }

// std::max<Scalar>(a, b) = (a < b) ? b : a; arap.h:199-218 pass the clamp first, so NaN
// yields the clamp.
__device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }

// ---- cotan weights (arap.h:182-239) -----------------------------------------------------

// One thread per face: the three clamped half-cotangents of arap.h:199-218, 225-227.
__global__ __launch_bounds__(kBlock) void k_cotan_face(LaunchSet L) {
#pragma clang fp contract(off)
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
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
    const MD_G MeshDev &M = L.meshes[k];
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
        M.w32[g] = (float)w;
        M.a64[g] = ax; M.a64[M.slots + g] = ay; M.a64[2 * M.slots + g] = az;
        M.a32[g] = (float)ax; M.a32[M.slots + g] = (float)ay; M.a32[2 * M.slots + g] = (float)az;
    }
}

// ---- local step (arap.h:354-384) ----------------------------------------------------------

// One Jacobi rotation of the symmetric 4x4 matrix a on the pair (P, Q); v accumulates the
// eigenvectors (columns).
template <typename T, int P, int Q>
__host__ __device__ __forceinline__ void jrot(T (&a)[4][4], T (&v)[4][4]) {
    T apq = a[P][Q];
    T scale = fabs(a[P][P]) + fabs(a[Q][Q]);
    if (!(fabs(apq) > (sizeof(T) == 4 ? T(1e-9) : T(1e-18)) * scale) ) {
        if (fabs(apq) <= (sizeof(T) == 4 ? T(1e-30) : T(1e-300))) return;
        if (scale > T(0)) return;
    }
    T theta = (a[Q][Q] - a[P][P]) * jrcp<T>(T(2) * apq);
    T t = (theta >= T(0) ? T(1) : T(-1)) * jrcp<T>(fabs(theta) + jsqrt<T>(theta * theta + T(1)));
    T c = jrsqrt<T>(t * t + T(1));
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
// Horn's symmetric 4x4 matrix N (J. Opt. Soc. Am. A 4(4), 1987).
//
// best_rotation_jacobi: cyclic Jacobi sweeps that start from the basis of the previous
// rotation (p = previous quaternion): the left-multiplication matrix L(p) is orthogonal with
// first column p, so L^T N L is already nearly diagonal when the rotation changed little and
// one or two sweeps converge. The converged eigenvector does not depend on the start.
template <typename T>
__host__ __device__ __forceinline__ void best_rotation_jacobi(const T (&m)[9], T px, T py, T pz, T pw, T &qx, T &qy,
                                                              T &qz, T &qw) {
    const T Sxx = m[0], Sxy = m[1], Sxz = m[2], Syx = m[3], Syy = m[4], Syz = m[5], Szx = m[6],
            Szy = m[7], Szz = m[8];
    const T N[4][4] = {{Sxx + Syy + Szz, Syz - Szy, Szx - Sxz, Sxy - Syx},
                       {Syz - Szy, Sxx - Syy - Szz, Sxy + Syx, Szx + Sxz},
                       {Szx - Sxz, Sxy + Syx, -Sxx + Syy - Szz, Syz + Szy},
                       {Sxy - Syx, Szx + Sxz, Syz + Szy, -Sxx - Syy + Szz}};
    {
        T pn = px * px + py * py + pz * pz + pw * pw;
        T inv = jrsqrt<T>(pn > T(0) ? pn : T(1));
        if (!(pn > T(0))) { px = 0; py = 0; pz = 0; pw = 1; }
        px *= inv; py *= inv; pz *= inv; pw *= inv;
    }
    // L(p) for p = (w, x, y, z) in the (scalar, vector) order of N's rows
    T v[4][4] = {{pw, -px, -py, -pz}, {px, pw, -pz, py}, {py, pz, pw, -px}, {pz, -py, px, pw}};
    T a[4][4];
    {
        T t[4][4];  // t = N L
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) t[r][c] = N[r][0] * v[0][c] + N[r][1] * v[1][c] + N[r][2] * v[2][c] + N[r][3] * v[3][c];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = r; c < 4; ++c) {
                T x = v[0][r] * t[0][c] + v[1][r] * t[1][c] + v[2][r] * t[2][c] + v[3][r] * t[3][c];
                a[r][c] = x;
                a[c][r] = x;
            }
    }
    constexpr int kSweeps = sizeof(T) == 4 ? 6 : 10;
    for (int sw = 0; sw < kSweeps; ++sw) {
        T off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[0][3]) + fabs(a[1][2]) + fabs(a[1][3]) + fabs(a[2][3]);
        T dia = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]) + fabs(a[3][3]);
        if (!(off > (sizeof(T) == 4 ? T(1e-8) : T(1e-17)) * dia)) break;
        jrot<T, 0, 1>(a, v); jrot<T, 0, 2>(a, v); jrot<T, 0, 3>(a, v);
        jrot<T, 1, 2>(a, v); jrot<T, 1, 3>(a, v); jrot<T, 2, 3>(a, v);
    }
    // column of the largest eigenvalue, by selects (an index into v would put v in scratch)
    T best = a[0][0];
    T q0 = v[0][0], q1 = v[1][0], q2 = v[2][0], q3 = v[3][0];
#pragma unroll
    for (int c = 1; c < 4; ++c) {
        const bool b = a[c][c] > best;
        best = b ? a[c][c] : best;
        q0 = b ? v[0][c] : q0;
        q1 = b ? v[1][c] : q1;
        q2 = b ? v[2][c] : q2;
        q3 = b ? v[3][c] : q3;
    }
    T nrm = jrsqrt<T>(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    // canonical sign: scalar part >= 0
    if (q0 < T(0)) nrm = -nrm;
    qw = q0 * nrm; qx = q1 * nrm; qy = q2 * nrm; qz = q3 * nrm;
}

// Cofactor (-1)^(I+J) det(a without row I and column J) of a 4x4 matrix (indices fixed at
// compile time, so a stays in registers).
template <typename T, int I, int J>
__host__ __device__ __forceinline__ T cofactor(const T (&a)[4][4]) {
    constexpr int r0 = I == 0 ? 1 : 0, r1 = I <= 1 ? 2 : 1, r2 = I <= 2 ? 3 : 2;
    constexpr int k0 = J == 0 ? 1 : 0, k1 = J <= 1 ? 2 : 1, k2 = J <= 2 ? 3 : 2;
    const T d = a[r0][k0] * (a[r1][k1] * a[r2][k2] - a[r1][k2] * a[r2][k1]) -
                a[r0][k1] * (a[r1][k0] * a[r2][k2] - a[r1][k2] * a[r2][k0]) +
                a[r0][k2] * (a[r1][k0] * a[r2][k1] - a[r1][k1] * a[r2][k0]);
    return ((I + J) & 1) ? -d : d;
}

// Column col of adj(a) for a symmetric 4x4 a.
template <typename T>
__host__ __device__ __forceinline__ void adj_column(const T (&a)[4][4], int col, T &v0, T &v1, T &v2, T &v3) {
    if (col == 0) {
        v0 = cofactor<T, 0, 0>(a); v1 = cofactor<T, 1, 0>(a); v2 = cofactor<T, 2, 0>(a); v3 = cofactor<T, 3, 0>(a);
    } else if (col == 1) {
        v0 = cofactor<T, 0, 1>(a); v1 = cofactor<T, 1, 1>(a); v2 = cofactor<T, 2, 1>(a); v3 = cofactor<T, 3, 1>(a);
    } else if (col == 2) {
        v0 = cofactor<T, 0, 2>(a); v1 = cofactor<T, 1, 2>(a); v2 = cofactor<T, 2, 2>(a); v3 = cofactor<T, 3, 2>(a);
    } else {
        v0 = cofactor<T, 0, 3>(a); v1 = cofactor<T, 1, 3>(a); v2 = cofactor<T, 2, 3>(a); v3 = cofactor<T, 3, 3>(a);
    }
}

// best_rotation: the largest eigenvalue lambda of N by Newton's method on its characteristic
// polynomial P(x) = x^4 + c2 x^2 + c1 x + c0 with c2 = -2 |M|_F^2, c1 = -8 det M, c0 = det N
// (N is traceless; Theobald, Acta Cryst. A61 (2005) 478-480, "quaternion characteristic
// polynomial"), started at sqrt(3/2 |c2|) = sqrt(3) |M|_F, an upper bound of lambda (the sum
// of M's singular values), from where Newton's iterates decrease monotonically to the largest
// root of a polynomial with real roots. The eigenvector is then the largest column of
// adj(N - lambda I) = prod_{j>1}(lambda_j - lambda) v v^T, taken at its largest diagonal entry.
// When that entry is small against |N|^3 the top eigenvalue is (nearly) repeated -- a rank-1
// or reflection-ambiguous covariance -- the adjugate loses accuracy, and the warm-started
// Jacobi sweeps decide instead. A zero covariance keeps the previous rotation.
template <typename T>
__host__ __device__ __forceinline__ bool best_rotation(const T (&m)[9], T px, T py, T pz, T pw, T &qx, T &qy, T &qz,
                                                       T &qw) {
    const T Sxx = m[0], Sxy = m[1], Sxz = m[2], Syx = m[3], Syy = m[4], Syz = m[5], Szx = m[6],
            Szy = m[7], Szz = m[8];
    const T f2 = Sxx * Sxx + Sxy * Sxy + Sxz * Sxz + Syx * Syx + Syy * Syy + Syz * Syz + Szx * Szx + Szy * Szy +
                 Szz * Szz;
    const T n00 = Sxx + Syy + Szz, n01 = Syz - Szy, n02 = Szx - Sxz, n03 = Sxy - Syx;
    const T n11 = Sxx - Syy - Szz, n12 = Sxy + Syx, n13 = Szx + Sxz;
    const T n22 = -Sxx + Syy - Szz, n23 = Syz + Szy;
    const T n33 = -Sxx - Syy + Szz;
    const T c2 = T(-2) * f2;
    const T c1 = T(-8) * (Sxx * (Syy * Szz - Syz * Szy) - Sxy * (Syx * Szz - Syz * Szx) + Sxz * (Syx * Szy - Syy * Szx));
    // det N from the 2x2 minors of rows (0, 1) and rows (2, 3)
    const T s0 = n00 * n11 - n01 * n01, s1 = n00 * n12 - n01 * n02, s2 = n00 * n13 - n01 * n03;
    const T s3 = n01 * n12 - n11 * n02, s4 = n01 * n13 - n11 * n03, s5 = n02 * n13 - n12 * n03;
    const T t5 = n22 * n33 - n23 * n23, t4 = n12 * n33 - n13 * n23, t3 = n12 * n23 - n13 * n22;
    const T t2 = n02 * n33 - n03 * n23, t1 = n02 * n23 - n03 * n22, t0 = n02 * n13 - n03 * n12;
    const T c0 = s0 * t5 - s1 * t4 + s2 * t3 + s3 * t2 - s4 * t1 + s5 * t0;
    if (!(f2 > T(0))) {  // no edges carry weight: keep the previous rotation
        T pn = px * px + py * py + pz * pz + pw * pw;
        T inv = jrsqrt<T>(pn > T(0) ? pn : T(1));
        if (!(pn > T(0))) { px = 0; py = 0; pz = 0; pw = 1; inv = 1; }
        if (pw < T(0)) inv = -inv;
        qx = px * inv; qy = py * inv; qz = pz * inv; qw = pw * inv;
        return true;
    }
    const T scale = jsqrt<T>(T(1.5) * (-c2));
    T lam = scale;
    constexpr T kTol = sizeof(T) == 4 ? T(1e-6) : T(1e-14);
    for (int it = 0; it < 60; ++it) {
        const T l2 = lam * lam;
        const T p = (l2 + c2) * l2 + c1 * lam + c0;
        const T dp = T(4) * l2 * lam + T(2) * c2 * lam + c1;
        if (!(dp > T(0))) break;  // at (or beyond) a multiple root: the fallback decides
        const T step = p / dp;
        lam -= step;
        if (!(fabs(step) > kTol * scale)) break;
    }
    // a = N - lambda I; the diagonal of adj(a) (principal 3x3 minors) picks the column
    const T a[4][4] = {{n00 - lam, n01, n02, n03},
                       {n01, n11 - lam, n12, n13},
                       {n02, n12, n22 - lam, n23},
                       {n03, n13, n23, n33 - lam}};
    const T d0 = cofactor<T, 0, 0>(a), d1 = cofactor<T, 1, 1>(a), d2 = cofactor<T, 2, 2>(a),
            d3 = cofactor<T, 3, 3>(a);
    T best = fabs(d0);
    int col = 0;
    if (fabs(d1) > best) { best = fabs(d1); col = 1; }
    if (fabs(d2) > best) { best = fabs(d2); col = 2; }
    if (fabs(d3) > best) { best = fabs(d3); col = 3; }
    constexpr T kGap = sizeof(T) == 4 ? T(1e-2) : T(1e-4);
    if (!(best > kGap * scale * scale * scale)) {
        best_rotation_jacobi<T>(m, px, py, pz, pw, qx, qy, qz, qw);
        return false;
    }
    // column `col` of adj(a) (a is symmetric, so adj(a) is too); then once more at the
    // Rayleigh quotient of that vector, whose error is the square of the vector's: the rounding
    // of det N moves the polynomial's root by up to eps |N|^4 / prod(gaps), and the vector's
    // error is that shift over the gap
    T v0, v1, v2, v3;
    adj_column<T>(a, col, v0, v1, v2, v3);
    {
        const T nv = v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3;
        const T w0 = n00 * v0 + n01 * v1 + n02 * v2 + n03 * v3, w1 = n01 * v0 + n11 * v1 + n12 * v2 + n13 * v3;
        const T w2 = n02 * v0 + n12 * v1 + n22 * v2 + n23 * v3, w3 = n03 * v0 + n13 * v1 + n23 * v2 + n33 * v3;
        const T rq = (v0 * w0 + v1 * w1 + v2 * w2 + v3 * w3) / nv;
        const T b[4][4] = {{n00 - rq, n01, n02, n03},
                           {n01, n11 - rq, n12, n13},
                           {n02, n12, n22 - rq, n23},
                           {n03, n13, n23, n33 - rq}};
        adj_column<T>(b, col, v0, v1, v2, v3);
    }
    T nrm = jrsqrt<T>(v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3);
    if (v0 < T(0)) nrm = -nrm;  // canonical sign: scalar part >= 0
    qw = v0 * nrm; qx = v1 * nrm; qy = v2 * nrm; qz = v3 * nrm;
    return true;
}

template <typename T>
__host__ __device__ __forceinline__ void quat_to_mat(T qx, T qy, T qz, T qw, T (&R)[9]) {
    T s = T(2) / (qx * qx + qy * qy + qz * qz + qw * qw);
    R[0] = T(1) - s * (qy * qy + qz * qz); R[1] = s * (qx * qy - qw * qz); R[2] = s * (qx * qz + qw * qy);
    R[3] = s * (qx * qy + qw * qz); R[4] = T(1) - s * (qx * qx + qz * qz); R[5] = s * (qy * qz - qw * qx);
    R[6] = s * (qx * qz - qw * qy); R[7] = s * (qy * qz + qw * qx); R[8] = T(1) - s * (qx * qx + qy * qy);
}

template <typename T> __device__ __forceinline__ const T *aplanes(const MD_G MeshDev &M);
template <> __device__ __forceinline__ const float *aplanes<float>(const MD_G MeshDev &M) { return M.a32; }
template <> __device__ __forceinline__ const double *aplanes<double>(const MD_G MeshDev &M) { return M.a64; }

// One lane per vertex (all n, constrained ones too, as arap.h:358 loops over every vertex):
// covariance sum_j w_ij (p_i - p_j)(p'_i - p'_j)^T (arap.h:366-374) from the slice's slots,
// then the rotation. The deformed edge p'_i - p'_j is formed in FP64 and rounded to T.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_local(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.n) return;
    const T *A = aplanes<T>(M);
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    d4 pi = M.pp[iv];
    T c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // 8 slots at a time: ids and edges first, then the neighbour gathers, then arithmetic
    // (slots past the slice width are clamped; they are padding with zero edges anyway)
    constexpr int kU = 8;
    for (int q0 = 0; q0 < wd; q0 += kU) {
        int jj[kU];
        T ex_[kU], ey_[kU], ez_[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            long long g = (long long)(s0 + min(q0 + u, wd - 1)) * kSlice + lane;
            jj[u] = M.sell_col[g];
            bool ok = q0 + u < wd;
            ex_[u] = ok ? A[g] : T(0);
            ey_[u] = ok ? A[M.slots + g] : T(0);
            ez_[u] = ok ? A[2 * M.slots + g] : T(0);
        }
        d4 pj[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) pj[u] = M.pp[jj[u]];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            T ax = ex_[u], ay = ey_[u], az = ez_[u];
            T ex = (T)(pi.x - pj[u].x), ey = (T)(pi.y - pj[u].y), ez = (T)(pi.z - pj[u].z);
            c[0] += ax * ex; c[1] += ax * ey; c[2] += ax * ez;
            c[3] += ay * ex; c[4] += ay * ey; c[5] += ay * ez;
            c[6] += az * ex; c[7] += az * ey; c[8] += az * ez;
        }
    }
    T qx, qy, qz, qw;
    V prev = reinterpret_cast<const V *>(M.quat)[iv];
    best_rotation<T>(c, prev.x, prev.y, prev.z, prev.w, qx, qy, qz, qw);
    V out;
    out.x = qx; out.y = qy; out.z = qz; out.w = qw;
    reinterpret_cast<V *>(M.quat)[iv] = out;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_identity_quat(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
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
// r_i of one free vertex, slots kU at a time (ids and edges, then the neighbour gathers).
template <typename T, int kU>
__device__ __forceinline__ typename V4T<T>::type residual_of(const MD_G MeshDev &M, int iv) {
    using V = typename V4T<T>::type;
    const int slice = iv / kSlice, lane = iv % kSlice;
    const T *A = aplanes<T>(M);
    const V *Q = reinterpret_cast<const V *>(M.quat);
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    d4 pi = M.pp[iv];
    V qi = Q[iv];
    T sax = 0, say = 0, saz = 0;   // sum_j a_ij
    T rx = 0, ry = 0, rz = 0;      // sum_j R_j a_ij
    double lx = 0, ly = 0, lz = 0; // sum_j w_ij (p'_i - p'_j)
    for (int q0 = 0; q0 < wd; q0 += kU) {
        int jj[kU];
        T ax_[kU], ay_[kU], az_[kU];
        double w_[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            long long g = (long long)(s0 + min(q0 + u, wd - 1)) * kSlice + lane;
            jj[u] = M.sell_col[g];
            bool ok = q0 + u < wd;
            ax_[u] = ok ? A[g] : T(0);
            ay_[u] = ok ? A[M.slots + g] : T(0);
            az_[u] = ok ? A[2 * M.slots + g] : T(0);
            w_[u] = ok ? M.w64[g] : 0.0;
        }
        V qj_[kU];
        d4 pj_[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) { qj_[u] = Q[jj[u]]; pj_[u] = M.pp[jj[u]]; }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            T ax = ax_[u], ay = ay_[u], az = az_[u];
            T Rj[9];
            quat_to_mat<T>(qj_[u].x, qj_[u].y, qj_[u].z, qj_[u].w, Rj);
            sax += ax; say += ay; saz += az;
            rx += Rj[0] * ax + Rj[1] * ay + Rj[2] * az;
            ry += Rj[3] * ax + Rj[4] * ay + Rj[5] * az;
            rz += Rj[6] * ax + Rj[7] * ay + Rj[8] * az;
            double w = w_[u];
            lx += w * (pi.x - pj_[u].x); ly += w * (pi.y - pj_[u].y); lz += w * (pi.z - pj_[u].z);
        }
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
    return out;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices_free) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.nf) return;
    reinterpret_cast<V *>(M.r)[iv] = residual_of<T, 8>(M, iv);
}

// ---- block LDL^T sweeps (arap.h:416-429) --------------------------------------------------

template <typename T> __device__ __forceinline__ const T *tptr(const void *p) {
    return reinterpret_cast<const T *>(p);
}

// Sum of cnt partial vectors U[k * stride + e], k = 0 .. cnt-1, in that order. One partial is
// one load; more are loaded 8 at a time (clamped addresses, masked values) so their
// latencies overlap.
constexpr int kPartialBatch = 8;
// The same in the panel-preloading tiles (few waves: registers to spare), so the 9-16 partials
// of a wide owner's x, or the 9-16 update slots of a wide front's position, arrive in one round.
constexpr int kWidePartialBatch = 16;
// Level launches of at most this many tiles load their panel entries ahead of the vector
// (latency-bound: few waves); larger ones keep the register count low for occupancy.
constexpr int kPreloadMaxTiles = 2048;

template <typename V, int B = kPartialBatch>
__device__ __forceinline__ V sum_partials(const V *U, int cnt, int stride, int e) {
    using T = decltype(V{}.x);
    if (cnt == 1) return U[e];
    V a;
    a.x = 0; a.y = 0; a.z = 0; a.w = 0;
    for (int k0 = 0; k0 < cnt; k0 += B) {
        V b[B];
#pragma unroll
        for (int u = 0; u < B; ++u) b[u] = U[(long long)min(k0 + u, cnt - 1) * stride + e];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            T m = (k0 + u < cnt) ? T(1) : T(0);
            a.x += m * b[u].x; a.y += m * b[u].y; a.z += m * b[u].z;
        }
    }
    return a;
}

// Children's partial updates at position pos of a level front: the sum of its uk slots (slot k
// at k * m + pos; the children's partials fill them in child order, unused ones are zero), so
// their addresses follow from the tile alone.
template <typename V, int B = kPartialBatch>
__device__ __forceinline__ V pull_slots(const MD_G MeshDev &M, const SnDev &S, int pos) {
    return sum_partials<V, B>(reinterpret_cast<const V *>(M.upart) + S.ubase, S.uk, S.w + S.noff, pos);
}

// The update-slot sums of two front positions (a: ona lanes, b: onb lanes), their loads
// issued together, B slots of each per round: the first column tile of a forward tile needs
// both its columns' and its off rows' sums, and one round trip instead of two brings them. The
// sums run in slot order, as pull_slots' do (inactive lanes read position 0 and add nothing).
template <typename V, int B>
__device__ __forceinline__ void pull_slots_pair(const MD_G MeshDev &M, const SnDev &S, int pa, bool ona, int pb,
                                                bool onb, V &a, V &b) {
    using T = decltype(V{}.x);
    const MD_G V *U = reinterpret_cast<const MD_G V *>(M.upart) + S.ubase;
    const int cnt = S.uk, stride = S.w + S.noff;
    const int ea = ona ? pa : 0, eb = onb ? pb : 0;
    a.x = 0; a.y = 0; a.z = 0; a.w = 0;
    b = a;
    if (cnt == 1) {
        const V ua = U[ea], ub = U[eb];
        if (ona) a = ua;
        if (onb) b = ub;
        return;
    }
    for (int k0 = 0; k0 < cnt; k0 += B) {
        V ua[B], ub[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const long long row = (long long)min(k0 + u, cnt - 1) * stride;
            ua[u] = U[row + ea];
            ub[u] = U[row + eb];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const T m = (k0 + u < cnt) ? T(1) : T(0);
            const T ma = ona ? m : T(0), mb = onb ? m : T(0);
            a.x += ma * ua[u].x; a.y += ma * ua[u].y; a.z += ma * ua[u].z;
            b.x += mb * ub[u].x; b.y += mb * ub[u].y; b.z += mb * ub[u].z;
        }
    }
}

// g[i] = P[min(i, n - 1) * stride] for i < N (n >= 1): a column of a panel tile loaded into
// registers ahead of the vector it multiplies; entries past n are valid panel values that meet
// zero vector entries.
template <int N, typename T>
__device__ __forceinline__ void panel_load(const T *P, long long stride, int n, T *g) {
#pragma unroll
    for (int i = 0; i < N; ++i) g[i] = P[(long long)min(i, n - 1) * stride];
}

// acc += sum_{i < N} g[i] * v[i] (v[i] = 0 for i >= n).
template <int N, typename T, typename V>
__device__ __forceinline__ void panel_dot(const T *g, const V *v, T &ax, T &ay, T &az) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const V x = v[i];
        ax += g[i] * x.x; ay += g[i] * x.y; az += g[i] * x.z;
    }
}

// acc += sum_{i < n} P[i * stride] * v[i] for n <= N (N = 16, 32 or 64): all N loads are
// issued together (addresses clamped to [0, n-1], or to 0 when n <= 0; values masked), so
// P[0] and v[0] must be readable.
template <int N, typename T, typename V>
__device__ __forceinline__ void tile_dot(const T *P, long long stride, const V *v, int n, T &ax, T &ay, T &az) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int ii = max(0, min(i, n - 1));
        T g = P[(long long)ii * stride];
        g = i < n ? g : T(0);
        V x = v[ii];
        ax += g * x.x; ay += g * x.y; az += g * x.z;
    }
}

template <typename T, typename V>
__device__ __forceinline__ void tile_dot_any(const T *P, long long stride, const V *v, int n, T &ax, T &ay, T &az) {
    if (n <= 16) tile_dot<16>(P, stride, v, n, ax, ay, az);
    else if (n <= 32) tile_dot<32>(P, stride, v, n, ax, ay, az);
    else tile_dot<64>(P, stride, v, n, ax, ay, az);
}

// Forward tile (s, r0, q0), one wavefront:
//   v_d[q0 .. q0+64) = r[c0 + q] + children's updates landing on those columns,
//   v_o[r0 .. r0+64) = children's updates landing on those off rows (first column tile),
//   z_s[q0 ..] = v_d (first row tile),
//   partial q0/64 of u_s on rows [r0, r0+64) = (q0 == 0 ? v_o : 0) - G_s[rows, cols] v_d.
// Children's updates arrive as kc partial vectors; each lane pulls the contributions listed
// for its position (host-built lists in child order), so no two lanes write the same place.
// fwd_tile loads the tile's G entries (N = 16, 32 or 64 columns) while the contributions arrive:
// a shorter chain of dependent loads for twice the registers (launches of few tiles);
// fwd_tile_lean loads them after the contributions.
template <int N, typename T>
__device__ __forceinline__ void fwd_tile(const MD_G MeshDev &M, const SnDev &S, int r0, int q0, int qn, int rn) {
    using V = typename V4T<T>::type;
    __shared__ V vd[kTile];
    __shared__ V vo[kTile];
    const int lane = threadIdx.x, w = S.w, no = S.noff, kcb = q0 / kTile;
    const V *R = reinterpret_cast<const V *>(M.r);
    V b, o;
    b.x = 0; b.y = 0; b.z = 0; b.w = 0;
    o = b;
    if (lane < qn) b = R[S.c0 + q0 + lane];
    // column-major G_s: entry (row r, column q) at q * no + r
    T g[N];
    if (rn > 0) panel_load<N, T>(tptr<T>(M.pc) + S.pc_off + (long long)q0 * no + r0 + min(lane, rn - 1), no, qn, g);
    const long long dst = lane < rn ? M.udst[S.roff + r0 + lane] : 0;  // this row's slot in the parent
    if (kcb == 0) {
        // columns and off rows: their slots in one round trip
        V c;
        pull_slots_pair<V, kPartialBatch>(M, S, q0 + lane, lane < qn, w + r0 + lane, lane < rn, c, o);
        if (lane < qn) { b.x += c.x; b.y += c.y; b.z += c.z; b.w = 0; }
    } else if (lane < qn) {
        const V c = pull_slots<V, kWidePartialBatch>(M, S, q0 + lane);
        b.x += c.x; b.y += c.y; b.z += c.z; b.w = 0;
    }
    vd[lane] = b;
    vo[lane] = o;
    __syncthreads();
    if (r0 == 0 && lane < qn) reinterpret_cast<V *>(M.z)[S.c0 + q0 + lane] = vd[lane];
    if (rn == 0) return;
    T ax = 0, ay = 0, az = 0;
    panel_dot<N, T, V>(g, vd, ax, ay, az);
    if (lane < rn) {
        V out;
        out.x = (kcb == 0 ? o.x : T(0)) - ax;
        out.y = (kcb == 0 ? o.y : T(0)) - ay;
        out.z = (kcb == 0 ? o.z : T(0)) - az;
        out.w = 0;
        reinterpret_cast<V *>(M.upart)[dst + (long long)kcb * S.dstride] = out;
    }
}

template <typename T>
__device__ __forceinline__ void fwd_tile_lean(const MD_G MeshDev &M, const SnDev &S, int r0, int q0, int qn, int rn) {
    using V = typename V4T<T>::type;
    __shared__ V vd[kTile];
    __shared__ V vo[kTile];
    const int lane = threadIdx.x, w = S.w, no = S.noff, kcb = q0 / kTile;
    const V *R = reinterpret_cast<const V *>(M.r);
    {
        V b;
        b.x = 0; b.y = 0; b.z = 0; b.w = 0;
        if (lane < qn) {
            b = R[S.c0 + q0 + lane];
            V c = pull_slots<V>(M, S, q0 + lane);
            b.x += c.x; b.y += c.y; b.z += c.z; b.w = 0;
        }
        vd[lane] = b;
        V o;
        o.x = 0; o.y = 0; o.z = 0; o.w = 0;
        if (kcb == 0 && lane < rn) o = pull_slots<V>(M, S, w + r0 + lane);
        vo[lane] = o;
    }
    __syncthreads();
    if (r0 == 0 && lane < qn) reinterpret_cast<V *>(M.z)[S.c0 + q0 + lane] = vd[lane];
    if (rn == 0) return;
    const T *G = tptr<T>(M.pc) + S.pc_off + (long long)q0 * no + r0 + min(lane, rn - 1);
    T ax = 0, ay = 0, az = 0;
    tile_dot_any<T, V>(G, no, vd, qn, ax, ay, az);
    if (lane < rn) {
        V o = vo[lane], out;
        out.x = (kcb == 0 ? o.x : T(0)) - ax;
        out.y = (kcb == 0 ? o.y : T(0)) - ay;
        out.z = (kcb == 0 ? o.z : T(0)) - az;
        out.w = 0;
        reinterpret_cast<V *>(M.upart)[M.udst[S.roff + r0 + lane] + (long long)kcb * S.dstride] = out;
    }
}

template <typename T, bool Pre>
__global__ __launch_bounds__(kTile) void k_fwd(LaunchSet L, int level) {
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const TileDev tl = L.tiles ? L.tiles[blockIdx.x] : M.ftile[L.tbase[k] + (blockIdx.x - L.prefix[k])];
    const SnDev &S = tl.s;
    const int r0 = tl.a, q0 = tl.b;
    const int qn = min(kTile, S.w - q0), rn = max(0, min(kTile, S.noff - r0));
    if (!Pre) fwd_tile_lean<T>(M, S, r0, q0, qn, rn);
    else if (qn <= 16) fwd_tile<16, T>(M, S, r0, q0, qn, rn);
    else if (qn <= 32) fwd_tile<32, T>(M, S, r0, q0, qn, rn);
    else fwd_tile<64, T>(M, S, r0, q0, qn, rn);
}

// Backward tile (s, q0, p0), one wavefront: partial p0/64 of x_s on columns [q0, q0+64):
//   sum over rows p in [p0, p0+64) of [Dinv_s; G_s][p, q] * t[p],  t = [z_s; -x_off],
// read from the row-major copy (lanes on consecutive columns). x at off rows is the sum of
// the owner's partials (xloc gives where they are). bwd_tile loads the tile's panel rows (N =
// 16, 32 or 64) while t is gathered; bwd_tile_lean loads them afterwards.
template <int N, typename T>
__device__ __forceinline__ void bwd_tile(const MD_G MeshDev &M, const SnDev &S, int q0, int p0, int qn, int pn) {
    using V = typename V4T<T>::type;
    __shared__ V tv[kTile];
    const int lane = threadIdx.x, w = S.w;
    const V *X = reinterpret_cast<const V *>(M.xpart);
    const int p = p0 + lane;
    // a root (no off rows) has no forward tile: its right-hand side and children's
    // contributions are gathered here (z = r + contributions, as the forward tile forms it)
    const bool root = S.noff == 0;
    V t;
    t.x = 0; t.y = 0; t.z = 0; t.w = 0;
    int4 xl = make_int4(0, 0, 0, 0);
    if (lane < pn) {
        if (p >= w) xl = M.xloc[S.roff + p - w];
        else if (root) t = reinterpret_cast<const V *>(M.r)[S.c0 + p];
        else t = reinterpret_cast<const V *>(M.z)[S.c0 + p];
    }
    T g[N];
    panel_load<N, T>(tptr<T>(M.pr) + S.pr_off + (long long)p0 * w + q0 + min(lane, qn - 1), w, pn, g);
    if (lane < pn && p >= w) {
        const V a = sum_partials<V, kWidePartialBatch>(X + xl.x, xl.z, xl.y, 0);
        t.x = -a.x; t.y = -a.y; t.z = -a.z;
    }
    if (lane < pn && p < w && root) {
        const V c = pull_slots<V, kWidePartialBatch>(M, S, p);
        t.x += c.x; t.y += c.y; t.z += c.z;
    }
    t.w = 0;
    tv[lane] = t;
    __syncthreads();
    T ax = 0, ay = 0, az = 0;
    panel_dot<N, T, V>(g, tv, ax, ay, az);
    if (lane < qn) {
        V out;
        out.x = ax; out.y = ay; out.z = az; out.w = 0;
        reinterpret_cast<V *>(M.xpart)[S.xoff + (long long)(p0 / kTile) * w + q0 + lane] = out;
    }
}

template <typename T>
__device__ __forceinline__ void bwd_tile_lean(const MD_G MeshDev &M, const SnDev &S, int q0, int p0, int qn, int pn) {
    using V = typename V4T<T>::type;
    __shared__ V tv[kTile];
    const int lane = threadIdx.x, w = S.w;
    const V *X = reinterpret_cast<const V *>(M.xpart);
    {
        V t;
        t.x = 0; t.y = 0; t.z = 0;
        int p = p0 + lane;
        if (lane < pn) {
            if (p < w && S.noff == 0) {  // a root: no forward tile (as in bwd_tile)
                t = reinterpret_cast<const V *>(M.r)[S.c0 + p];
                const V c = pull_slots<V>(M, S, p);
                t.x += c.x; t.y += c.y; t.z += c.z;
            } else if (p < w) {
                t = reinterpret_cast<const V *>(M.z)[S.c0 + p];
            } else {
                int4 xl = M.xloc[S.roff + p - w];
                V a = sum_partials<V>(X + xl.x, xl.z, xl.y, 0);
                t.x = -a.x; t.y = -a.y; t.z = -a.z;
            }
        }
        t.w = 0;
        tv[lane] = t;
    }
    __syncthreads();
    const T *P = tptr<T>(M.pr) + S.pr_off + (long long)p0 * w + q0 + min(lane, qn - 1);
    T ax = 0, ay = 0, az = 0;
    tile_dot_any<T, V>(P, w, tv, pn, ax, ay, az);
    if (lane < qn) {
        V out;
        out.x = ax; out.y = ay; out.z = az; out.w = 0;
        reinterpret_cast<V *>(M.xpart)[S.xoff + (long long)(p0 / kTile) * w + q0 + lane] = out;
    }
}

template <typename T, bool Pre>
__global__ __launch_bounds__(kTile) void k_bwd(LaunchSet L, int level) {
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const TileDev tl = L.tiles ? L.tiles[blockIdx.x] : M.btile[L.tbase[k] + (blockIdx.x - L.prefix[k])];
    const SnDev &S = tl.s;
    const int q0 = tl.a, p0 = tl.b;
    const int qn = min(kTile, S.w - q0), pn = min(kTile, S.w + S.noff - p0);
    if (!Pre) bwd_tile_lean<T>(M, S, q0, p0, qn, pn);
    else if (pn <= 16) bwd_tile<16, T>(M, S, q0, p0, qn, pn);
    else if (pn <= 32) bwd_tile<32, T>(M, S, q0, p0, qn, pn);
    else bwd_tile<64, T>(M, S, q0, p0, qn, pn);
}

// acc += sum_{i < n} P[i * stride] * v[i] (n <= 64; n <= 0 adds nothing but still reads P[0]
// and v[0]) with the unrolled width chosen per wavefront (the widest active lane's class,
// 16 / 32 / 64), so lanes of different widths do not diverge.
template <typename T, typename V>
__device__ __forceinline__ void wave_dot(const T *P, long long stride, const V *v, int n, T &ax, T &ay, T &az) {
    if (__any(n > 32)) tile_dot<64>(P, stride, v, n, ax, ay, az);
    else if (__any(n > 16)) tile_dot<32>(P, stride, v, n, ax, ay, az);
    else tile_dot<16>(P, stride, v, n, ax, ay, az);
}

// Record of one supernode of a bottom subtree (16 ints, as the host lays it out).
struct SubNode {
    int c0, w, noff, fb;     // columns, off rows, forward front base in LDS
    int pc, pr, cp, ob;      // G offset, [Dinv; G] offset, contribution entries, off-row sources
    int uo, xo, xs, pa;      // root: first update-slot entry (udst) or -1; x output, LDS column base
                             //   of x, position prefix
    int pb, pq, fbb, pad;    // off-row prefix, column prefix, backward front base in LDS
};

// Last record j in [a, b) whose prefix (field F) is <= t.
template <int SubNode::*F>
__device__ __forceinline__ int sub_find(const SubNode *N, int a, int b, int t) {
    int lo = a, hi = b - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (N[mid].*F <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Copies the records [sd.x, sd.y) of a subtree to LDS.
__device__ __forceinline__ void sub_records(const MD_G MeshDev &M, const int4 &sd, SubNode *N) {
    int4 *dst = reinterpret_cast<int4 *>(N);
    for (int i = threadIdx.x; i < 4 * (sd.y - sd.x); i += kTile) dst[i] = M.snode[4 * sd.x + i];
    __syncthreads();
}

// Bottom subtree, forward sweep, one wavefront per subtree (sd = first record, end record,
// height table offset, heights). Heights bottom-up, each in two phases with every front in
// LDS:
//   positions: front[p] = (p < w ? r[c0 + p] : 0) + children's update entries; z = front[p < w],
//   off rows:  front[w + i] -= G_s[i, 0..w) front[0 .. w)  (u_s in place; the root also stores
//              u_s for its parent outside the subtree).
// With Fused (one-mesh launches), the residual is formed here: subtree rows straight into
// their front positions, level rows by blocks past the subtrees (no k_residual launch; more
// registers, so batches keep the separate residual).
template <typename T, bool Fused>
__global__ __launch_bounds__(kTile) void k_sub_fwd(LaunchSet L, int node_cap) {
    using V = typename V4T<T>::type;
    extern __shared__ __align__(32) unsigned char sub_smem[];
    SubNode *N = reinterpret_cast<SubNode *>(sub_smem);
    V *F = reinterpret_cast<V *>(sub_smem + sizeof(SubNode) * node_cap);
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const int lane = threadIdx.x;
    const int bk = blockIdx.x - L.prefix[k];
    if (Fused && bk >= M.nsub) {
        // blocks past the subtrees: the residual of 64 rows of level supernodes
        const int j = (bk - M.nsub) * kTile + lane;
        if (j < M.nlrows) {
            const int i = M.lrows[j];
            reinterpret_cast<V *>(M.r)[i] = residual_of<T, 8>(M, i);
        }
        return;
    }
    const int4 sd = M.sub[bk];
    const V *R = reinterpret_cast<const V *>(M.r);
    V *Z = reinterpret_cast<V *>(M.z);
    V *U = reinterpret_cast<V *>(M.upart);
    const T *G = tptr<T>(M.pc);
    if (Fused) {
        // the residual of the subtree's rows, straight into their front positions (it is
        // read by nothing else)
        const int2 rr = M.srow[bk];
        for (int t = lane; t < rr.y; t += kTile) F[M.sfi[rr.x + t]] = residual_of<T, 8>(M, rr.x + t);
    }
    sub_records(M, sd, N);  // ends with a barrier
    for (int h = 0; h < sd.w; ++h) {
        const int a = M.sub_h[sd.z + h] - sd.x, b = M.sub_h[sd.z + h + 1] - sd.x;
        const int npos = N[b - 1].pa + N[b - 1].w + N[b - 1].noff;
        for (int t = lane; t < npos; t += kTile) {
            const SubNode &S = N[sub_find<&SubNode::pa>(N, a, b, t)];
            const int p = t - S.pa;
            V v;
            v.x = 0; v.y = 0; v.z = 0; v.w = 0;
            if (p < S.w) v = Fused ? F[S.fb + p] : R[S.c0 + p];  // the residual
            // children's update entries: two inline, the rest from a -1-terminated list
            const int2 e = M.scp[S.cp + p];
            if (e.x >= 0) {
                const V u = F[e.x];
                v.x += u.x; v.y += u.y; v.z += u.z;
            }
            if (e.y >= 0) {
                const V u = F[e.y];
                v.x += u.x; v.y += u.y; v.z += u.z;
            } else if (e.y <= -2) {
                for (int c = -2 - e.y;; ++c) {
                    const int q = M.sfc[c];
                    if (q < 0) break;
                    const V u = F[q];
                    v.x += u.x; v.y += u.y; v.z += u.z;
                }
            }
            v.w = 0;
            F[S.fb + p] = v;
            if (p < S.w) Z[S.c0 + p] = v;
        }
        __syncthreads();
        const int nrow = N[b - 1].pb + N[b - 1].noff;
        for (int t0 = 0; t0 < nrow; t0 += kTile) {
            const int t = t0 + lane;
            const bool on = t < nrow;
            const SubNode &S = N[sub_find<&SubNode::pb>(N, a, b, on ? t : nrow - 1)];
            const int i = (on ? t : nrow - 1) - S.pb;
            T ax = 0, ay = 0, az = 0;
            // column-major G_s: entry (row i, column q) at q * noff + i
            wave_dot<T, V>(G + S.pc + i, S.noff, F + S.fb, on ? S.w : 0, ax, ay, az);
            if (on) {
                V o = F[S.fb + S.w + i];
                o.x -= ax; o.y -= ay; o.z -= az;
                F[S.fb + S.w + i] = o;
                if (S.uo >= 0) U[M.udst[S.uo + i]] = o;  // the root: into the parent's slot
            }
        }
        __syncthreads();
    }
}

// Bottom subtree, backward sweep, one wavefront per subtree. Heights top-down, with the
// subtree's solution columns and the fronts of one height in LDS:
//   positions: t[p] = z[c0 + p] (p < w), else -x of the off row, from LDS when the row lies
//              in the subtree, else the sum of its owner's partials,
//   columns:   x_s[q] = sum_p [Dinv_s; G_s][p, q] t[p], to LDS and to the single partial.
template <typename T>
__global__ __launch_bounds__(kTile) void k_sub_bwd(LaunchSet L, int node_cap) {
    using V = typename V4T<T>::type;
    extern __shared__ __align__(32) unsigned char sub_smem[];
    SubNode *N = reinterpret_cast<SubNode *>(sub_smem);
    V *F = reinterpret_cast<V *>(sub_smem + sizeof(SubNode) * node_cap);
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const int lane = threadIdx.x;
    const V *X = reinterpret_cast<const V *>(M.xpart);
    const int bk = blockIdx.x - L.prefix[k];
    if (bk >= M.nsub) {
        // blocks past the subtrees: p' += x for 64 rows of level supernodes (x summed from
        // the partials of the level sweeps)
        const int j = (bk - M.nsub) * kTile + lane;
        if (j < M.nlrows) {
            const int i = M.lrows[j];
            const int4 ri = M.rowinfo[i];
            const V a = sum_partials<V>(X + ri.x, ri.z, ri.y, 0);
            d4 pv = M.pp[i];
            pv.x += (double)a.x; pv.y += (double)a.y; pv.z += (double)a.z;
            M.pp[i] = pv;
        }
        return;
    }
    const int4 sd = M.sub[bk];
    const T *P = tptr<T>(M.pr);
    // the column region F[0 .. ncols) starts with z of the subtree's rows (read at the row's
    // height, then replaced by its x)
    const int2 rr = M.srow[bk];
    {
        const V *Zs = reinterpret_cast<const V *>(M.z);
        for (int t = lane; t < rr.y; t += kTile) F[t] = Zs[rr.x + t];
    }
    sub_records(M, sd, N);  // ends with a barrier
    for (int h = sd.w - 1; h >= 0; --h) {
        const int a = M.sub_h[sd.z + h] - sd.x, b = M.sub_h[sd.z + h + 1] - sd.x;
        const int npos = N[b - 1].pa + N[b - 1].w + N[b - 1].noff;
        for (int t = lane; t < npos; t += kTile) {
            const SubNode &S = N[sub_find<&SubNode::pa>(N, a, b, t)];
            const int p = t - S.pa;
            V v;
            if (p < S.w) {
                v = F[S.xs + p];
            } else {
                const int2 d = M.sbo[S.ob + p - S.w];
                // x of the row: from LDS, or the owner's partials loaded 8 at a time
                const V g = d.y < 0 ? F[d.x] : sum_partials<V>(X + d.x, d.y >> 16, d.y & 0xffff, 0);
                v.x = -g.x; v.y = -g.y; v.z = -g.z;
            }
            v.w = 0;
            F[S.fbb + p] = v;
        }
        __syncthreads();
        const int ncol = N[b - 1].pq + N[b - 1].w;
        for (int t0 = 0; t0 < ncol; t0 += kTile) {
            const int t = t0 + lane;
            const bool on = t < ncol;
            const SubNode &S = N[sub_find<&SubNode::pq>(N, a, b, on ? t : ncol - 1)];
            const int q = (on ? t : ncol - 1) - S.pq, m = S.w + S.noff;
            T ax = 0, ay = 0, az = 0;
            // row-major [Dinv_s; G_s]: entry (row p, column q) at p * w + q
            wave_dot<T, V>(P + S.pr + q, S.w, F + S.fbb, on ? min(m, kTile) : 0, ax, ay, az);
            if (__any(on && m > kTile)) {
                // rows 64 .. m-1; lanes with m <= 64 keep their own base and add nothing
                const bool two = on && m > kTile;
                wave_dot<T, V>(P + S.pr + q + (two ? (long long)kTile * S.w : 0), S.w, F + S.fbb + (two ? kTile : 0),
                               two ? m - kTile : 0, ax, ay, az);
            }
            if (on) {
                V o;
                o.x = ax; o.y = ay; o.z = az; o.w = 0;
                F[S.xs + q] = o;
            }
        }
        __syncthreads();
    }
    // x of the subtree's rows is final and read by nothing outside the subtree: p' += x
    for (int t = lane; t < rr.y; t += kTile) {
        const V o = F[t];
        d4 pv = M.pp[rr.x + t];
        pv.x += (double)o.x; pv.y += (double)o.y; pv.z += (double)o.z;
        M.pp[rr.x + t] = pv;
    }
}

// p'_i += x_i for every free vertex, x_i the sum of its supernode's partials.
// ---- Jacobi-preconditioned CG on the correction system (alternative to the sweeps) -------
//
// Solves A d = r for the three coordinates at once (three independent CG recurrences, one
// per coordinate), A = L of arap.h:325-332 over the free vertices, M = diag(A). d starts at 0:
// the residual r = b - A p' already warm-starts from the previous iterate. Every block works
// on 256 consecutive free vertices; dot products are per-block partials that each block of
// the next kernel reduces in the same fixed order (deterministic, no atomics).

template <typename T> __device__ __forceinline__ const T *wts(const MD_G MeshDev &M);
template <> __device__ __forceinline__ const float *wts<float>(const MD_G MeshDev &M) { return M.w32; }
template <> __device__ __forceinline__ const double *wts<double>(const MD_G MeshDev &M) { return M.w64; }

// Two block sums at once (six values in one tree): a into pa[blk], b into pb[blk].
template <typename T, typename V>
__device__ __forceinline__ void block_partial2(const T (&a)[3], const T (&b)[3], MD_G V *pa, MD_G V *pb, int blk,
                                               T (*red)[kBlock]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        red[q][threadIdx.x] = a[q];
        red[3 + q][threadIdx.x] = b[q];
    }
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        V va, vb;
        va.x = red[0][0]; va.y = red[1][0]; va.z = red[2][0]; va.w = 0;
        vb.x = red[3][0]; vb.y = red[4][0]; vb.z = red[5][0]; vb.w = 0;
        pa[blk] = va;
        pb[blk] = vb;
    }
}

// Two partial arrays summed at once (fixed order); the results broadcast through LDS.
template <typename T, typename V>
__device__ __forceinline__ void reduce_partials2(const MD_G V *pa, const MD_G V *pb, int nb, T (&oa)[3], T (&ob)[3],
                                                 T (*red)[kBlock]) {
    T a[6] = {0, 0, 0, 0, 0, 0};
    for (int k = threadIdx.x; k < nb; k += kBlock) {
        const V x = pa[k], y = pb[k];
        a[0] += x.x; a[1] += x.y; a[2] += x.z;
        a[3] += y.x; a[4] += y.y; a[5] += y.z;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
#pragma unroll
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        oa[q] = red[q][0];
        ob[q] = red[3 + q][0];
    }
    __syncthreads();
}

// L_ii = sum of every neighbour weight (arap.h:332), once per factorisation.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_diag(LaunchSet L) {
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices_free) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.nf) return;
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    double d = 0.0;
    for (int q = 0; q < wd; ++q) d += M.w64[(long long)(s0 + q) * kSlice + lane];
    reinterpret_cast<T *>(M.diag)[iv] = (T)d;
}

// Jacobi-PCG in the Chronopoulos-Gear form (Chronopoulos and Gear, J. Comput. Appl. Math. 25
// (1989) 153-168): with u = M^-1 r and w = A u, step i forms
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M^-1 r, w = A u,
//   gamma = (r, u), delta = (w, u),
// and the next step takes beta = gamma_new / gamma, alpha = gamma_new / (delta - beta gamma_new /
// alpha): one reduction, so one launch, per CG iteration. The new u of a neighbour is formed from
// its r, w and s of the previous step, which live in the other buffer of each pair (parity i),
// so no lane reads what another writes. Per coordinate, a solve whose gamma fell below
// tol^2 gamma_0 keeps alpha = 0 (its x and r stay).
template <typename T, typename V> struct CgVecs {
    MD_G V *r[2], *w[2], *s[2], *x, *p;
};
template <typename T, typename V>
__device__ __forceinline__ CgVecs<T, V> cg_vecs(const MD_G MeshDev &M) {
    CgVecs<T, V> c;
    MD_G V *v = reinterpret_cast<MD_G V *>(M.cg_v);
    const long long nf = M.nf;
    c.r[0] = reinterpret_cast<MD_G V *>(M.r);
    c.r[1] = v;
    c.w[0] = v + nf;
    c.w[1] = v + 2 * nf;
    c.s[0] = v + 3 * nf;
    c.s[1] = v + 4 * nf;
    c.x = reinterpret_cast<MD_G V *>(M.cg_x);
    c.p = reinterpret_cast<MD_G V *>(M.cg_p);
    return c;
}

// x = p = s_0 = 0, u = M^-1 r, w_0 = A u, partials of (r, u) and (w, u) (parity 0). One lane
// per free vertex in slice order (block b: vertices [256 b, 256 b + 256)).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(LaunchSet L) {
    using V = typename V4T<T>::type;
    __shared__ T red[6][kBlock];
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const int blk = blockIdx.x - L.prefix[k];
    const int slice = blk * kSlicesPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int iv = slice * kSlice + lane;
    const CgVecs<T, V> c = cg_vecs<T, V>(M);
    const MD_G T *D = reinterpret_cast<const MD_G T *>(M.diag);
    T gam[3] = {0, 0, 0}, del[3] = {0, 0, 0};
    if (slice < M.nslices_free && iv < M.nf) {
        const T *W = wts<T>(M);
        const int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
        const V r = c.r[0][iv];
        const T di = D[iv], rdi = T(1) / di;
        const T ux = r.x * rdi, uy = r.y * rdi, uz = r.z * rdi;
        T yx = di * ux, yy = di * uy, yz = di * uz;
        for (int q = 0; q < wd; ++q) {
            const long long g = (long long)(s0 + q) * kSlice + lane;
            const int j = M.sell_col[g];
            if (j != iv && j < M.nf) {
                const V rj = c.r[0][j];
                const T wq = W[g] / D[j];
                yx -= wq * rj.x; yy -= wq * rj.y; yz -= wq * rj.z;
            }
        }
        V z;
        z.x = 0; z.y = 0; z.z = 0; z.w = 0;
        V w;
        w.x = yx; w.y = yy; w.z = yz; w.w = 0;
        c.w[0][iv] = w;
        c.s[0][iv] = z;
        c.x[iv] = z;
        c.p[iv] = z;
        gam[0] = r.x * ux; gam[1] = r.y * uy; gam[2] = r.z * uz;
        del[0] = yx * ux; del[1] = yy * uy; del[2] = yz * uz;
    }
    MD_G V *part = reinterpret_cast<MD_G V *>(M.cg_part);
    const int nb = M.nblk_free;
    block_partial2<T, V>(gam, del, part, part + nb, blk, red);
}

// Step i (parity c = i & 1): reads r, w, s and the partials of parity c, writes those of 1 - c.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_step(LaunchSet L, int step) {
    using V = typename V4T<T>::type;
    __shared__ T red[6][kBlock];
    const int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    const int blk = blockIdx.x - L.prefix[k];
    const int slice = blk * kSlicesPerBlock + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int iv = slice * kSlice + lane;
    const int c = step & 1, n = 1 - c;
    const CgVecs<T, V> v = cg_vecs<T, V>(M);
    const MD_G T *D = reinterpret_cast<const MD_G T *>(M.diag);
    MD_G V *part = reinterpret_cast<MD_G V *>(M.cg_part);
    const int nb = M.nblk_free;
    T ga[3], de[3];
    reduce_partials2<T, V>(part + 2 * c * nb, part + (2 * c + 1) * nb, nb, ga, de, red);
    MD_G V *scal = part + 4 * nb;  // (gamma, alpha) of each parity, then the first gamma
    T al[3], be[3];
    const V gp = scal[2 * c], ap = scal[2 * c + 1], gz = scal[4];
    const T gprev[3] = {gp.x, gp.y, gp.z}, aprev[3] = {ap.x, ap.y, ap.z};
    const T g0[3] = {step == 0 ? ga[0] : gz.x, step == 0 ? ga[1] : gz.y, step == 0 ? ga[2] : gz.z};
    for (int q = 0; q < 3; ++q) {
        const bool conv = (double)ga[q] <= M.pcg_tol2 * (double)g0[q];
        if (step == 0) {
            be[q] = T(0);
            al[q] = (!conv && de[q] != T(0)) ? ga[q] / de[q] : T(0);
        } else if (conv || aprev[q] == T(0) || gprev[q] == T(0)) {
            be[q] = T(0);
            al[q] = T(0);
        } else {
            be[q] = ga[q] / gprev[q];
            const T den = de[q] - be[q] * ga[q] / aprev[q];
            al[q] = den != T(0) ? ga[q] / den : T(0);
        }
    }
    if (blockIdx.x == L.prefix[k] && threadIdx.x == 0) {  // for the next step
        V g, a;
        g.x = ga[0]; g.y = ga[1]; g.z = ga[2]; g.w = 0;
        a.x = al[0]; a.y = al[1]; a.z = al[2]; a.w = 0;
        scal[2 * n] = g;
        scal[2 * n + 1] = a;
        if (step == 0) scal[4] = g;  // the first gamma, for the stopping test
    }
    T gam[3] = {0, 0, 0}, del[3] = {0, 0, 0};
    if (slice < M.nslices_free && iv < M.nf) {
        const T *W = wts<T>(M);
        const int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
        const T di = D[iv], rdi = T(1) / di;
        const V r = v.r[c][iv], w = v.w[c][iv], so = v.s[c][iv];
        V p = v.p[iv], x = v.x[iv];
        // p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s
        p.x = r.x * rdi + be[0] * p.x; p.y = r.y * rdi + be[1] * p.y; p.z = r.z * rdi + be[2] * p.z;
        V sn;
        sn.x = w.x + be[0] * so.x; sn.y = w.y + be[1] * so.y; sn.z = w.z + be[2] * so.z; sn.w = 0;
        x.x += al[0] * p.x; x.y += al[1] * p.y; x.z += al[2] * p.z;
        V rn;
        rn.x = r.x - al[0] * sn.x; rn.y = r.y - al[1] * sn.y; rn.z = r.z - al[2] * sn.z; rn.w = 0;
        v.p[iv] = p;
        v.x[iv] = x;
        v.s[n][iv] = sn;
        v.r[n][iv] = rn;
        // w = A u with u = M^-1 r of the new r; a neighbour's new u from its previous r, w, s
        const T ux = rn.x * rdi, uy = rn.y * rdi, uz = rn.z * rdi;
        T yx = di * ux, yy = di * uy, yz = di * uz;
        for (int q = 0; q < wd; ++q) {
            const long long g = (long long)(s0 + q) * kSlice + lane;
            const int j = M.sell_col[g];
            if (j != iv && j < M.nf) {
                const V rj = v.r[c][j], wj = v.w[c][j], sj = v.s[c][j];
                const T wq = W[g] / D[j];
                yx -= wq * (rj.x - al[0] * (wj.x + be[0] * sj.x));
                yy -= wq * (rj.y - al[1] * (wj.y + be[1] * sj.y));
                yz -= wq * (rj.z - al[2] * (wj.z + be[2] * sj.z));
            }
        }
        V wn;
        wn.x = yx; wn.y = yy; wn.z = yz; wn.w = 0;
        v.w[n][iv] = wn;
        gam[0] = rn.x * ux; gam[1] = rn.y * uy; gam[2] = rn.z * uz;
        del[0] = yx * ux; del[1] = yy * uy; del[2] = yz * uz;
    }
    block_partial2<T, V>(gam, del, part + 2 * n * nb, part + (2 * n + 1) * nb, blk, red);
}

// p' += d for the free vertices.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_apply(LaunchSet L) {
    using V = typename V4T<T>::type;
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int i = (blockIdx.x - L.prefix[k]) * kBlock + threadIdx.x;
    if (i >= M.nf) return;
    V x = reinterpret_cast<const V *>(M.cg_x)[i];
    d4 pv = M.pp[i];
    pv.x += (double)x.x; pv.y += (double)x.y; pv.z += (double)x.z;
    M.pp[i] = pv;
}

// ---- vertex normals of the deformed mesh (for rendering clients) --------------------------

// Area-weighted face normal of every face from p' (cross product of two edges, unnormalised),
// three doubles per face into fn.
__global__ __launch_bounds__(kBlock) void k_face_normals(LaunchSet L, double *fn) {
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int f = (blockIdx.x - L.prefix[k]) * kBlock + threadIdx.x;
    if (f >= M.F) return;
    d4 a = M.pp[M.tri[3 * f]], b = M.pp[M.tri[3 * f + 1]], c = M.pp[M.tri[3 * f + 2]];
    double ux = b.x - a.x, uy = b.y - a.y, uz = b.z - a.z;
    double vx = c.x - a.x, vy = c.y - a.y, vz = c.z - a.z;
    fn[3 * f] = uy * vz - uz * vy;
    fn[3 * f + 1] = uz * vx - ux * vz;
    fn[3 * f + 2] = ux * vy - uy * vx;
}

// Per vertex: every incident face appears on exactly two of its one-ring slots (the two face
// edges at the vertex), so half the sum over the slots' face lists is the sum of incident
// face normals; normalised (zero stays zero). Output in internal order.
__global__ __launch_bounds__(kBlock) void k_vertex_normals(LaunchSet L, const double *fn, d4 *out) {
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
    int slice = (blockIdx.x - L.prefix[k]) * kSlicesPerBlock + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (slice >= M.nslices) return;
    int iv = slice * kSlice + lane;
    if (iv >= M.n) return;
    int s0 = M.slice_ptr[slice], wd = M.slice_ptr[slice + 1] - s0;
    double x = 0, y = 0, z = 0;
    for (int q = 0; q < wd; ++q) {
        long long g = (long long)(s0 + q) * kSlice + lane;
        for (int e = M.slot_fe_ptr[g]; e < M.slot_fe_ptr[g + 1]; ++e) {
            int f = M.slot_fe[e] / 3;
            x += fn[3 * f]; y += fn[3 * f + 1]; z += fn[3 * f + 2];
        }
    }
    double nrm = sqrt(x * x + y * y + z * z);
    double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    d4 o;
    o.x = x * inv; o.y = y * inv; o.z = z * inv; o.w = 0.0;
    out[iv] = o;
}

// Internal-order vectors (positions or normals) to the caller's packed xyz in original vertex
// order: dst[3 perm[iv] + c] = v[iv].c (perm = original id of each internal id, a bijection).
__global__ __launch_bounds__(kBlock) void k_pack_xyz(const d4 *nrm, const int *perm, int n, double *dst) {
    int iv = blockIdx.x * kBlock + threadIdx.x;
    if (iv >= n) return;
    d4 v = nrm[iv];
    double *o = dst + 3 * (size_t)perm[iv];
    o[0] = v.x; o[1] = v.y; o[2] = v.z;
}

// Device workspace of the entry points that allocate nothing (md_mesh_normals_noalloc): one
// array per GPU, set aside when the code object loads and held until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress returns on each device (the
// array's name in host code is a host-side stand-in, not a device address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// ---- ARAP energy -------------------------------------------------------------------------

// E = sum_i sum_j w_ij |(p'_i - p'_j) - R_i (p_i - p_j)|^2 over all vertices, in FP64; one
// partial sum per block of 256 vertices, reduced in a fixed order.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_energy(LaunchSet L) {
    using V = typename V4T<T>::type;
    __shared__ double red[kBlock];
    int k = find_mesh(L, blockIdx.x);
    const MD_G MeshDev &M = L.meshes[k];
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

// p'[iv[k]] = xyz[k] for k targets (internal ids), used to move handles without a re-setup.
// The caller's packed xyz (original order) into internal order: p[iv] = xyz[perm[iv]] (p may
// be null), and pp[iv] the same for free vertices, the constrained ones (iv >= nf) taking their
// targets tgt[iv - nf] (internal order) instead (pp may be null).
__global__ __launch_bounds__(kBlock) void k_unpack_xyz(const double *xyz, const int *perm, int n, int nf,
                                                       const double *tgt, d4 *p, d4 *pp) {
    const int iv = blockIdx.x * kBlock + threadIdx.x;
    if (iv >= n) return;
    const double *s = xyz + 3 * (size_t)perm[iv];
    const d4 v{s[0], s[1], s[2], 0.0};
    if (p) p[iv] = v;
    if (pp) {
        if (iv < nf) pp[iv] = v;
        else {
            const double *t = tgt + 3 * (size_t)(iv - nf);
            pp[iv] = d4{t[0], t[1], t[2], 0.0};
        }
    }
}

// Targets of several meshes: entry t writes its position into the p' entry at its address.
__global__ __launch_bounds__(kBlock) void k_scatter_entries(const TargetEntry *e, int k) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= k) return;
    const TargetEntry x = e[t];
    *reinterpret_cast<MD_G d4 *>(x.addr) = d4{x.x, x.y, x.z, 0.0};
}

// ---- numeric factorization (arap.h:337, SimplicialLDLT::compute) ------------------------
// Multifrontal block LDL^T in FP64, one height of the supernodal tree at a time, one
// workgroup per supernode. Every supernode's front lives in its own workspace blocks; the
// children's update matrices are added in child order, so the factor is deterministic.

constexpr int kFacBlock = 256;
constexpr int kAsmU = 8;      // child update entries per assembly thread in flight
constexpr int kAsmRows = 2048;  // off rows of a front searched in LDS (longer lists: in memory)

// L_jj = sum of the weights of every neighbour of j, in slot order, and their maximum.
__global__ __launch_bounds__(kBlock) void k_fac_diag(FacArgs A) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= A.nf) return;
    const int sl = j / kSlice, l = j % kSlice;
    double acc = 0.0;
    for (int q = A.slice_ptr[sl]; q < A.slice_ptr[sl + 1]; ++q) {
        const long long g = (long long)q * kSlice + l;
        if (A.sell_col[g] != j) acc += A.w64[g];
    }
    A.diag[j] = acc;
    atomicMax(A.amax, (unsigned long long)__double_as_longlong(fabs(acc)));
}

// Front of each supernode of one height: the pivot block and G rows (panel), F_os (copy),
// the update matrix; original entries (L_jj and -w_ij for free i > j, slot order) first,
// then the children's update matrices in child order. `per` workgroups per supernode, each
// owning `cols` columns of the front (pivot columns: P and F; later ones: U), so no two write
// the same entry and every entry receives its additions in that fixed order.
constexpr int kAsmSlotLanes = 16;  // threads per pivot column for its original entries
constexpr int kAsmChildren = 64;   // children whose records are staged in LDS at a time
__global__ __launch_bounds__(kFacBlock) void k_fac_assemble(FacArgs A, int h0, int per, int cols) {
    const FacDev S = A.sn[A.hlist[h0 + blockIdx.x / per]];
    const int tid = threadIdx.x, w = S.w, no = S.noff, m = S.m;
    const int t0 = (int)(blockIdx.x % per) * cols, t1 = min(m, t0 + cols);
    if (t0 >= m) return;
    MD_G double *P = A.P + S.poff, *F = A.S + S.soff, *U = A.U + S.uoff;
    const MD_G int *rows = A.offrows + S.roff;
    // the front's row list in LDS for the searches of the original entries (pivot columns), and
    // each child's row map in turn
    __shared__ int rsh[kAsmRows];
    __shared__ int relsh[kAsmRows];
    // the children's (rows, row map, update matrix) records, kAsmChildren at a time (the first
    // batch loaded up front, beside the zeroing)
    __shared__ int chn[kAsmChildren], chr[kAsmChildren];
    __shared__ long long chu[kAsmChildren];
    const int nch = S.ce - S.cb;
    auto stage_children = [&](int c0) {
        if (tid < min(nch - c0, kAsmChildren)) {
            const FacDev C = A.sn[A.child[S.cb + c0 + tid]];
            chn[tid] = C.noff;
            chr[tid] = C.roff;
            chu[tid] = C.uoff;
        }
    };
    static_assert(kAsmChildren <= kFacBlock, "one child record per thread");
    stage_children(0);
    const bool staged = t0 < w && no <= kAsmRows;
    if (staged)
        for (int e = tid; e < no; e += kFacBlock) rsh[e] = rows[e];
    // zero the owned columns: contiguous ranges of P and F (pivot columns) and of U
    const int kend = min(t1, w);
    if (t0 < w) {
        for (long long e = (long long)t0 * m + tid; e < (long long)kend * m; e += kFacBlock) P[e] = 0.0;
        for (long long e = (long long)t0 * no + tid; e < (long long)kend * no; e += kFacBlock) F[e] = 0.0;
    }
    if (t1 > w)
        for (long long e = (long long)(max(t0, w) - w) * no + tid; e < (long long)(t1 - w) * no; e += kFacBlock)
            U[e] = 0.0;
    __syncthreads();
    // original entries, kAsmSlotLanes threads per pivot column (one slot each); each entry of the
    // front receives at most one, right after its zero, so it is stored
    for (int k = t0 + tid / kAsmSlotLanes; k < kend; k += kFacBlock / kAsmSlotLanes) {
        const int j = S.c0 + k, sl = j / kSlice, l = j % kSlice;
        if (tid % kAsmSlotLanes == 0) P[(long long)k * m + k] = A.diag[j];
        for (int q = A.slice_ptr[sl] + tid % kAsmSlotLanes; q < A.slice_ptr[sl + 1]; q += kAsmSlotLanes) {
            const long long g = (long long)q * kSlice + l;
            const int i = A.sell_col[g];
            if (i <= j || i >= A.nf) continue;
            const double v = -A.w64[g];
            if (i < S.c0 + w) {
                P[(long long)k * m + (i - S.c0)] = v;
            } else {
                int lo = 0, hi = no;  // first off row >= i
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((staged ? rsh[mid] : rows[mid]) < i) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < no && (staged ? rsh[lo] : rows[lo]) == i) F[(long long)k * no + lo] = v;
                else *A.flag = 2;
            }
        }
    }
    __syncthreads();
    // children's update matrices, child by child: the child's lower entries (a >= b) of its
    // columns b landing in [t0, t1) (rel is increasing, so they are the columns [blo, bend),
    // counted over the workgroup), flattened over the workgroup, kAsmU entries per thread in
    // flight (their loads, then the read-modify-writes: distinct entries of one child land on
    // distinct front entries)
    for (int c = 0; c < nch; ++c) {
        if (c > 0 && c % kAsmChildren == 0) {  // the previous child's loop ended with a barrier
            stage_children(c);
            __syncthreads();
        }
        const int nc = chn[c % kAsmChildren], roff = chr[c % kAsmChildren];
        const long long uoff = chu[c % kAsmChildren];
        const MD_G int *relg = A.rel + roff;
        const bool rstaged = nc <= kAsmRows;
        int blo = 0, bend = 0;
        for (int a0 = 0; a0 < nc; a0 += kFacBlock) {
            const int a = a0 + tid;
            const int r = a < nc ? relg[a] : m;
            if (rstaged && a < nc) relsh[a] = r;
            blo += __syncthreads_count(r < t0);
            bend += __syncthreads_count(r < t1);
        }
        const MD_G double *Uc = A.U + uoff;
        const int total = (bend - blo) * nc;
        for (int e0 = tid; e0 < total; e0 += kFacBlock * kAsmU) {
            double v[kAsmU];
            MD_G double *dst[kAsmU];
#pragma unroll
            for (int u = 0; u < kAsmU; ++u) {
                const int e = e0 + u * kFacBlock, bb = e / nc, a = e - bb * nc, b = blo + bb;
                dst[u] = nullptr;
                v[u] = 0.0;
                if (e < total && a >= b) {
                    const int ra = rstaged ? relsh[a] : relg[a], rb = rstaged ? relsh[b] : relg[b];
                    v[u] = Uc[(long long)b * nc + a];
                    if (ra >= m || rb > ra) *A.flag = 2;
                    else if (rb < w) dst[u] = ra < w ? P + (long long)rb * m + ra : F + (long long)rb * no + (ra - w);
                    else dst[u] = U + (long long)(rb - w) * no + (ra - w);
                }
            }
            double old[kAsmU];
#pragma unroll
            for (int u = 0; u < kAsmU; ++u) old[u] = dst[u] ? *dst[u] : 0.0;
#pragma unroll
            for (int u = 0; u < kAsmU; ++u)
                if (dst[u]) *dst[u] = old[u] + v[u];
        }
        __syncthreads();
    }
}

// The pivot block's Cholesky factor and its inverse X = L^{-1}, blocked by kFb columns
// (right-looking): for each column block s, the diagonal block is factored and inverted (the
// diagonal block of X, row-major in the T workspace; k_fac_potrf for block 0, later blocks by the
// workgroup of k_fac_syrk that completes them), k_fac_trsm forms the rows below it (times that
// inverse) and row block s of X, k_fac_syrk updates the trailing lower blocks and accumulates the
// X blocks below: X_IJ = -X_II W_IJ with W_IJ = sum_{K=J}^{I-1} L_IK X_KJ, the term K = s added at
// step s (in X's place), so W_sJ is complete when step s begins.
constexpr int kFb = 64;
// operands of the MFMA products are staged in LDS kMk values of k at a time, rows kMpad doubles apart
constexpr int kMk = 32;
constexpr int kMpad = 66;

// Diagonal block s, by four wavefronts (unrolled over the block's columns: a loop over them with
// the same work per step measured almost twice as slow). Cholesky (unblocked, right-looking) as an
// unscaled LDL^T: lane i of wave v holds row i of the block's columns [16 v, 16 v + 16) (a[j] =
// A(kb + i, kb + 16 v + j), then u). The wave owning a panel of kPanel columns factors it in
// registers: at step k it takes the pivot d_k = u_k(k) from lane k (v_readlane, two 32-bit halves
// per double), forms 1/d_k (v_rcp_f64 and one cubic Newton step) and updates the panel's later
// columns, a(i, c) -= (u_k(i) / d_k) u_k(c) (entries above the diagonal too, never used). It puts
// u_k and u_k / d_k in LDS; after one barrier the waves with columns past the panel apply the
// rank-kPanel update, the next panel's wave first and in the same basic block as that panel's
// factorisation, so the chain starts while the rest of its update is still being issued. At the
// end every wave scales its columns, L(:, k) = u_k / sqrt(d_k) (v_rsq_f64 and two Newton steps).
//
// Inverse X = L^{-1} by 16 x 16 blocks: wave v inverts diagonal block v by columns (four lanes per
// column, lane r of column i holding X(t, i) for t = r mod 4; X(k, i) = (delta_ki - sum_{t<k}
// L(k, t) X(t, i)) / L(k, k), the sum split over the four lanes and joined by a DPP butterfly), and
// after one barrier wave J forms block column J downwards on the matrix cores, X_IJ = -X_II W_IJ
// with W_IJ = sum_{K=J}^{I-1} L_IK X_KJ. X is zero above the diagonal, exactly. Rows and columns
// past a partial block (b < 64) hold zero in L and in the slot of 1 / L(k, k), so X is zero there.
typedef double dv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// v from the lane whose index is this lane's xor Q (Q = 1, 2: within a quad), by DPP.
template <int Q>
__device__ __forceinline__ double quad_xor(double v) {
    constexpr int ctrl = Q == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

constexpr int kPotrfCols = kFb / 4;  // columns per wave
constexpr int kLdsPad = kFb + 1;
constexpr int kIb = 16;              // blocks of the inverse (one per wave)
static_assert(kFb == 4 * kIb && kFacBlock == 4 * 64, "four wavefronts, one 16-row block of the inverse each");

// kPanel columns per barrier. FULL: a block of kFb columns (no per-column bound checks, so a
// panel's chain is one basic block); else b columns, the rest zero. Panel buffers: cu[parity][k][i]
// = u_k(i), cv[parity][k][i] = u_k(i) / d_k.
constexpr int kPanel = 8;
static_assert(kPotrfCols % kPanel == 0, "a panel lies within one wave's columns");
template <int K, bool FULL>
__device__ __forceinline__ void factor_panel(double (&a)[kPotrfCols], double (*cu)[kPanel][kFb], double (*cv)[kPanel][kFb],
                                             int lane, int b, double amax, bool &bad) {
    constexpr int j0 = K % kPotrfCols, buf = (K / kPanel) & 1;
#pragma unroll
    for (int kk = 0; kk < kPanel; ++kk) {
        double v = 0.0;  // columns past the block stay zero
        if (FULL || K + kk < b) {
            const double d = readlane_f64(a[j0 + kk], K + kk);
            if (!(d > 1e-14 * amax)) bad = true;
            // 1/d: v_rcp_f64 (relative error below 2^-22) and y (1 + e + e^2), e = 1 - d y
            double y = __builtin_amdgcn_rcp(d);
            const double e = __builtin_fma(-d, y, 1.0);
            y = __builtin_fma(y, __builtin_fma(e, e, e), y);
            v = a[j0 + kk] * y;
#pragma unroll
            for (int c = kk + 1; c < kPanel; ++c) a[j0 + c] -= v * readlane_f64(a[j0 + kk], K + c);
        }
        cu[buf][kk][lane] = a[j0 + kk];
        cv[buf][kk][lane] = v;
    }
}

// Panel K is in LDS once the barrier is passed: the update of the columns past it (from column jn of
// the next panel's wave, all columns of the waves after it), then the next panel's factorisation.
template <int K, bool FULL>
__device__ __forceinline__ void chol_step(double (&a)[kPotrfCols], double (*cu)[kPanel][kFb], double (*cv)[kPanel][kFb],
                                          int wv, int lane, int b, double amax, bool &bad) {
    if constexpr (K < kFb) {
        if (FULL || K < b) {
            constexpr int buf = (K / kPanel) & 1;
            __syncthreads();
            constexpr int nxt = K + kPanel, nown = nxt / kPotrfCols, jn = nxt % kPotrfCols;
            if constexpr (nxt < kFb) {
                double lik[kPanel];
#pragma unroll
                for (int kk = 0; kk < kPanel; ++kk) lik[kk] = cv[buf][kk][lane];
                if (wv == nown) {
#pragma unroll
                    for (int j = jn; j < kPotrfCols; ++j)
#pragma unroll
                        for (int kk = 0; kk < kPanel; ++kk) a[j] -= lik[kk] * cu[buf][kk][kPotrfCols * nown + j];
                    if (FULL || nxt < b) factor_panel<nxt, FULL>(a, cu, cv, lane, b, amax, bad);
                } else if (wv > nown) {
#pragma unroll
                    for (int j = 0; j < kPotrfCols; ++j)
#pragma unroll
                        for (int kk = 0; kk < kPanel; ++kk) a[j] -= lik[kk] * cu[buf][kk][kPotrfCols * wv + j];
                }
            }
            chol_step<nxt, FULL>(a, cu, cv, wv, lane, b, amax, bad);
        }
    }
}

// Diagonal block o / 16 of the inverse: lane (4 ii + r) forms X(o + t, o + ii) for t = r mod 4,
// x[q] for t = 4 q + r; the slot (k, k) of Lsh holds 1 / L(k, k).
template <int K>
__device__ __forceinline__ void inv_diag(const double (*Lsh)[kLdsPad], double (&x)[kIb / 4], int r, int ii, int o) {
    if constexpr (K < kIb) {
        double part = 0.0;
#pragma unroll
        for (int q = 0; q < (K + 3) / 4; ++q) {
            const int t = 4 * q + r;
            part += t < K ? Lsh[o + K][o + t] * x[q] : 0.0;
        }
        part += quad_xor<1>(part);
        part += quad_xor<2>(part);
        const double xk = ((ii == K ? 1.0 : 0.0) - part) * Lsh[o + K][o + K];
        if (r == (K & 3)) x[K / 4] = xk;
        inv_diag<K + 1>(Lsh, x, r, ii, o);
    }
}

// Block column J of the inverse on the matrix cores (v_mfma_f64_16x16x4f64: A(i, k) from lane
// i + 16 k, B(k, j) from lane j + 16 k, D(i, j) in lane j + 16 (i % 4), register i / 4). Register c
// of a D block is the B operand of k in [4 c, 4 c + 4), so the blocks X_KJ feed the next products
// as they are. Writes the block column to T (row-major), zero above the diagonal block.
template <int J>
__device__ __forceinline__ void inv_blockcol(const double (*Lsh)[kLdsPad], const double (*Xd)[kIb][kIb + 1], int lane,
                                             MD_G double *T, int w, int kb, int b) {
    const int i = lane & 15, m = lane >> 4;
    dv4 xb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) xb[J][c] = Xd[J][4 * c + m][i];
#pragma unroll
    for (int I = J + 1; I < 4; ++I) {
        dv4 acc = dv4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int K = J; K < I; ++K)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Lsh[kIb * I + i][kIb * K + 4 * c + m], xb[K][c], acc, 0, 0, 0);
        dv4 d = dv4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 4; ++c) d = __builtin_amdgcn_mfma_f64_16x16x4f64(Xd[I][i][4 * c + m], acc[c], d, 0, 0, 0);
        xb[I] = -d;
    }
    const int colx = kIb * J + i;
    if (colx < b) {
#pragma unroll
        for (int I = 0; I < 4; ++I)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int row = kIb * I + 4 * c + m;
                if (row < b) T[(long long)(kb + row) * w + kb + colx] = I < J ? 0.0 : xb[I][c];
            }
    }
}

// LDS of the pivot block, two arrays: the panel buffers (later the inverse's diagonal blocks; sized
// to also hold the staged A operand of k_fac_syrk), and L (also holds the staged B operand).
constexpr int kPbPanels = 2 * 2 * kPanel * kFb;
constexpr int kPbPan = kPbPanels > kMk * kMpad ? kPbPanels : kMk * kMpad;
constexpr int kPbL = kFb * kLdsPad > kMk * kMpad ? kFb * kLdsPad : kMk * kMpad;
static_assert(kFb / kIb * kIb * (kIb + 1) <= kPbPanels, "the diagonal blocks of the inverse fit the panel buffers");

// Diagonal block s of front S by the whole workgroup: L into the panel, X = L^{-1} into T. FROM_LDS:
// the block's lower entries are in lsh (row-major, rows kLdsPad apart), not yet in the panel.
template <bool FROM_LDS>
__device__ __forceinline__ void pivot_block(const FacArgs &A, const FacDev &S, int s, double *pan, double *lsh) {
    const int w = S.w, m = S.m, kb = s * kFb;
    if (kb >= w) return;
    const int b = min(kFb, w - kb), wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    MD_G double *P = A.P + S.poff;
    MD_G double *T = A.T + S.toff;
    double (*cu)[kPanel][kFb] = reinterpret_cast<double (*)[kPanel][kFb]>(pan);
    double (*cv)[kPanel][kFb] = cu + 2;
    double (*Lsh)[kLdsPad] = reinterpret_cast<double (*)[kLdsPad]>(lsh);
    double a[kPotrfCols];
#pragma unroll
    for (int j = 0; j < kPotrfCols; ++j) {
        // column-major panel: A(kb + lane, kb + c) at (kb + c) m + kb + lane (lower part only)
        const int c = kPotrfCols * wv + j;
        if constexpr (FROM_LDS) a[j] = (lane < b && c < b && c <= lane) ? lsh[lane * kLdsPad + c] : 0.0;
        else a[j] = (lane < b && c < b && c <= lane) ? P[(long long)(kb + c) * m + kb + lane] : 0.0;
    }
    if constexpr (FROM_LDS) __syncthreads();  // every entry read before lsh takes L
    const double amax = __longlong_as_double((long long)*A.amax);
    bool bad = false;
    if (b == kFb) {
        if (wv == 0) factor_panel<0, true>(a, cu, cv, lane, b, amax, bad);
        chol_step<0, true>(a, cu, cv, wv, lane, b, amax, bad);
    } else {
        if (wv == 0) factor_panel<0, false>(a, cu, cv, lane, b, amax, bad);
        chol_step<0, false>(a, cu, cv, wv, lane, b, amax, bad);
    }
    if (bad && threadIdx.x == 0) *A.flag = 1;
    // L(:, k) = u_k / sqrt(d_k), d_k = u_k(k); lower part to the panel, all of the wave's columns to
    // LDS (the diagonal as 1 / L(i, i))
    double rdiag = 0.0;
#pragma unroll
    for (int j = 0; j < kPotrfCols; ++j) {
        const int c = kPotrfCols * wv + j;
        const double d = readlane_f64(a[j], c);
        double rd = 0.0;
        if (c < b) {
            rd = __builtin_amdgcn_rsq(d);
            rd = rd * (1.5 - 0.5 * d * rd * rd);
            rd = rd * (1.5 - 0.5 * d * rd * rd);
        }
        if (lane == c) rdiag = rd;
        a[j] *= rd;
    }
#pragma unroll
    for (int j = 0; j < kPotrfCols; ++j) {
        const int c = kPotrfCols * wv + j;
        if (lane < b && c < b && c <= lane) P[(long long)(kb + c) * m + kb + lane] = a[j];
        Lsh[lane][c] = c == lane ? rdiag : a[j];
    }
    __syncthreads();
    double (*Xd)[kIb][kIb + 1] = reinterpret_cast<double (*)[kIb][kIb + 1]>(pan);
    {
        const int r = lane & 3, ii = lane >> 2;
        double x[kIb / 4] = {0.0, 0.0, 0.0, 0.0};
        inv_diag<0>(Lsh, x, r, ii, kIb * wv);
#pragma unroll
        for (int q = 0; q < kIb / 4; ++q) Xd[wv][4 * q + r][ii] = x[q];
    }
    __syncthreads();
    if (wv == 0) inv_blockcol<0>(Lsh, Xd, lane, T, w, kb, b);
    else if (wv == 1) inv_blockcol<1>(Lsh, Xd, lane, T, w, kb, b);
    else if (wv == 2) inv_blockcol<2>(Lsh, Xd, lane, T, w, kb, b);
    else inv_blockcol<3>(Lsh, Xd, lane, T, w, kb, b);
}

__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_fac_potrf(FacArgs A, int h0, int s) {
    __shared__ double pan[kPbPanels];
    __shared__ double lsh[kFb * kLdsPad];
    pivot_block<false>(A, A.sn[A.hlist[h0 + blockIdx.x]], s, pan, lsh);
}

// ---- FP64 dense products of the factorization on the matrix cores ----------------------
//
// v_mfma_f64_16x16x4f64 (one wavefront): D(16x16) = A(16x4) B(4x16) + C with A(i, k) from
// lane i + 16 k, B(k, j) from lane j + 16 k, and D(i, j) in lane j + 16 (i % 4), register
// i / 4. A 64 x 64 output tile per workgroup of 256 threads: wave v computes rows
// [32 (v / 2), +32) and columns [32 (v % 2), +32) as 2 x 2 such blocks, with the operands
// staged in LDS kMk values of k at a time (As[k][i] = A(i, k), Bs[k][j] = B(k, j)).

struct MmaAcc {
    dv4 c[2][2];
};

__device__ __forceinline__ void mma_zero(MmaAcc &acc) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc.c[x][y] = dv4{0.0, 0.0, 0.0, 0.0};
}

// acc += As^T Bs over the staged kMk values of k.
__device__ __forceinline__ void mma_chunk(const double (*As)[kMpad], const double (*Bs)[kMpad], MmaAcc &acc) {
    const int wv = threadIdx.x / 64, lane = threadIdx.x & 63;
    const int rb = 32 * (wv >> 1) + (lane & 15), cb = 32 * (wv & 1) + (lane & 15);
#pragma unroll
    for (int k4 = 0; k4 < kMk; k4 += 4) {
        const int kk = k4 + (lane >> 4);
        const double a0 = As[kk][rb], a1 = As[kk][rb + 16];
        const double b0 = Bs[kk][cb], b1 = Bs[kk][cb + 16];
        acc.c[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc.c[0][0], 0, 0, 0);
        acc.c[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc.c[0][1], 0, 0, 0);
        acc.c[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc.c[1][0], 0, 0, 0);
        acc.c[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc.c[1][1], 0, 0, 0);
    }
}

// Tile coordinates (i, j) within the 64 x 64 output of this thread's element (x, y, r).
__device__ __forceinline__ int mma_row(int x, int r) {
    return 32 * ((threadIdx.x / 64) >> 1) + 16 * x + 4 * r + ((threadIdx.x & 63) >> 4);
}
__device__ __forceinline__ int mma_col(int y) { return 32 * ((threadIdx.x / 64) & 1) + 16 * y + (threadIdx.x & 15); }

// acc += sum over nchunks staged chunks of kMk values of k, software-pipelined: each thread
// holds its kPer values of both operands of the next chunk in registers, loaded (fetch(c, ra,
// rb)) while the matrix cores work on the current one, and put into As / Bs (put(ra, rb)) once
// every wave is done reading them. Thread t stages the entries e = t + kFacBlock i, i < kPer.
constexpr int kPer = kMk * 64 / kFacBlock;
template <class Fetch, class Put>
__device__ __forceinline__ void mma_pipeline(int nchunks, Fetch fetch, Put put, double (*As)[kMpad],
                                             double (*Bs)[kMpad], MmaAcc &acc) {
    double ra[kPer], rb[kPer];
    if (nchunks > 0) fetch(0, ra, rb);
    for (int c = 0; c < nchunks; ++c) {
        __syncthreads();
        put(ra, rb);
        __syncthreads();
        if (c + 1 < nchunks) fetch(c + 1, ra, rb);
        mma_chunk(As, Bs, acc);
    }
}

// Lower block (I, J), I >= J, of the q-th tile of a lower-triangular tile grid.
__device__ __forceinline__ void lower_tile(int q, int &I, int &J) {
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= q) ++i;
    I = i;
    J = q - i * (i + 1) / 2;
}

// Split-K of the long products (few tiles, long K: the top heights): each of `ks` workgroups of a
// tile multiplies over its share of the K chunks and stores its accumulators (kAccVals doubles per
// workgroup, coalesced; MODE 1), and a second launch (MODE 2) sums the ks partials in split order
// (deterministic) and finishes the tile as the unsplit kernel (MODE 0) does. (A last-arriving
// workgroup summing in the same launch measured several times slower: every workgroup's
// device-scope fence writes back its L2.)
constexpr int kAccVals = 16 * kFacBlock;
static_assert(kAccVals == fac_split_k_doubles(1), "partials per workgroup as the host sizes them");
__device__ __forceinline__ void acc_store(MD_G double *wp, const MmaAcc &acc) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) wp[((x * 2 + y) * 4 + r) * kFacBlock + threadIdx.x] = acc.c[x][y][r];
}
// Sums in split order; the loads of four splits are issued before their additions.
__device__ __forceinline__ void acc_sum(const MD_G double *wp, int ks, MmaAcc &acc) {
    for (int z0 = 0; z0 < ks; z0 += 4) {
        double v[4][16];
#pragma unroll
        for (int dz = 0; dz < 4; ++dz)
#pragma unroll
            for (int t = 0; t < 16; ++t)
                v[dz][t] = z0 + dz < ks ? wp[(long long)(z0 + dz) * kAccVals + t * kFacBlock + threadIdx.x] : 0.0;
#pragma unroll
        for (int dz = 0; dz < 4; ++dz)
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc.c[x][y][r] += v[dz][(x * 2 + y) * 4 + r];
    }
}

// D^{-1} = X^T X into the pivot block of the panel (symmetric): lower 64 x 64 tile (I, J)
// of D^{-1}(r, c) = sum_{k >= r} X(k, r) X(k, c), X = L^{-1} row-major in T (zero above its
// diagonal inside the diagonal blocks), `per` workgroups per supernode.
template <int MODE>
__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 2 : 4))) void k_fac_dinv(FacArgs A, int h0, int per, int ks) {
    const int tile = MODE == 1 ? (int)blockIdx.x / ks : (int)blockIdx.x, z = MODE == 1 ? (int)blockIdx.x % ks : 0;
    MmaAcc acc;
    mma_zero(acc);
    // MODE 2: the partials' loads go out first (a tile with no work stored none and returns below)
    if constexpr (MODE == 2) acc_sum(A.Wk + (long long)tile * ks * kAccVals, ks, acc);
    const FacDev S = A.sn[A.hlist[h0 + tile / per]];
    const int w = S.w, m = S.m, nb = (w + 63) / 64, q = tile % per;
    if (q >= nb * (nb + 1) / 2) return;
    int I, J;
    lower_tile(q, I, J);
    const int r0 = 64 * I, c0 = 64 * J;
    double *P = A.P + S.poff;
    const double *X = A.T + S.toff;
    __shared__ double As[kMk][kMpad];
    __shared__ double Bs[kMk][kMpad];
    if constexpr (MODE != 2) {
        const int nch = (w - r0 + kMk - 1) / kMk, cb = z * nch / ks, ce = (z + 1) * nch / ks;
        mma_pipeline(
            ce - cb,
            [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i, x = e % 64, k = r0 + (cb + c) * kMk + e / 64;
                    const bool kin = k < w;
                    ra[i] = kin && r0 + x < w ? X[(long long)k * w + r0 + x] : 0.0;
                    rb[i] = kin && c0 + x < w ? X[(long long)k * w + c0 + x] : 0.0;
                }
            },
            [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i;
                    As[e / 64][e % 64] = ra[i];
                    Bs[e / 64][e % 64] = rb[i];
                }
            },
            As, Bs, acc);
    }
    if constexpr (MODE == 1) {
        acc_store(A.Wk + ((long long)tile * ks + z) * kAccVals, acc);
        return;
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = r0 + mma_row(x, r), cc = c0 + mma_col(y);
                if (rr < w && cc < w && cc <= rr) {
                    const double v = acc.c[x][y][r];
                    P[(long long)cc * m + rr] = v;
                    P[(long long)rr * m + cc] = v;
                    // the backward panel [D^{-1}; G] (row-major) in the solve precision
                    const long long o1 = S.pr_off + (long long)rr * w + cc, o2 = S.pr_off + (long long)cc * w + rr;
                    if (A.solve64) {
                        static_cast<MD_G double *>(A.pr)[o1] = v;
                        static_cast<MD_G double *>(A.pr)[o2] = v;
                    } else {
                        static_cast<MD_G float *>(A.pr)[o1] = (float)v;
                        static_cast<MD_G float *>(A.pr)[o2] = (float)v;
                    }
                }
            }
}

// G = F_os D^{-1} into panel rows [w, m): 64 x 64 tile (rows of G, columns), `per`
// workgroups per supernode. A(i, k) = F_os(r0 + i, k) (column-major), B(k, j) = D^{-1}(k, q0 + j).
template <int MODE>
__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 2 : 4))) void k_fac_g(FacArgs A, int h0, int per, int ks) {
    const int tile = MODE == 1 ? (int)blockIdx.x / ks : (int)blockIdx.x, z = MODE == 1 ? (int)blockIdx.x % ks : 0;
    MmaAcc acc;
    mma_zero(acc);
    // MODE 2: the partials' loads go out first (a tile with no work stored none and returns below)
    if constexpr (MODE == 2) acc_sum(A.Wk + (long long)tile * ks * kAccVals, ks, acc);
    const FacDev S = A.sn[A.hlist[h0 + tile / per]];
    const int w = S.w, no = S.noff, m = S.m, nr = (no + 63) / 64, nq = (w + 63) / 64;
    const int q = tile % per;
    if (q >= nr * nq) return;
    const int r0 = 64 * (q % nr), q0 = 64 * (q / nr);
    double *P = A.P + S.poff;
    const double *F = A.S + S.soff;
    __shared__ double As[kMk][kMpad];
    __shared__ double Bs[kMk][kMpad];
    if constexpr (MODE != 2) {
        const int nch = (w + kMk - 1) / kMk, cb = z * nch / ks, ce = (z + 1) * nch / ks;
        mma_pipeline(
            ce - cb,
            [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i, x = e % 64, k = (cb + c) * kMk + e / 64;
                    ra[i] = k < w && r0 + x < no ? F[(long long)k * no + r0 + x] : 0.0;
                    // B(k, q0 + x): consecutive threads read consecutive k of one column
                    const int xb = e / kMk, kg = (cb + c) * kMk + e % kMk;
                    rb[i] = kg < w && q0 + xb < w ? P[(long long)(q0 + xb) * m + kg] : 0.0;
                }
            },
            [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i;
                    As[e / 64][e % 64] = ra[i];
                    Bs[e % kMk][e / kMk] = rb[i];
                }
            },
            As, Bs, acc);
    }
    if constexpr (MODE == 1) {
        acc_store(A.Wk + ((long long)tile * ks + z) * kAccVals, acc);
        return;
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = r0 + mma_row(x, r), qq = q0 + mma_col(y);
                if (rr < no && qq < w) {
                    const double v = acc.c[x][y][r];
                    P[(long long)qq * m + w + rr] = v;
                    // the forward panel G (column-major) and the backward panel's G rows
                    const long long oc = S.pc_off + (long long)qq * no + rr, orr = S.pr_off + (long long)(w + rr) * w + qq;
                    if (A.solve64) {
                        static_cast<MD_G double *>(A.pc)[oc] = v;
                        static_cast<MD_G double *>(A.pr)[orr] = v;
                    } else {
                        static_cast<MD_G float *>(A.pc)[oc] = (float)v;
                        static_cast<MD_G float *>(A.pr)[orr] = (float)v;
                    }
                }
            }
}

// U -= G F_os^T (lower part): lower 64 x 64 tile (I, J) of U(a, b), a >= b, `per` workgroups
// per supernode. A(i, k) = G(a0 + i, k) (panel rows [w, m)), B(k, j) = F_os(b0 + j, k).
template <int MODE>
__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 2 : 4))) void k_fac_u(FacArgs A, int h0, int per, int ks) {
    const int tile = MODE == 1 ? (int)blockIdx.x / ks : (int)blockIdx.x, z = MODE == 1 ? (int)blockIdx.x % ks : 0;
    MmaAcc acc;
    mma_zero(acc);
    // MODE 2: the partials' loads go out first (a tile with no work stored none and returns below)
    if constexpr (MODE == 2) acc_sum(A.Wk + (long long)tile * ks * kAccVals, ks, acc);
    const FacDev S = A.sn[A.hlist[h0 + tile / per]];
    const int w = S.w, no = S.noff, m = S.m, nb = (no + 63) / 64, q = tile % per;
    if (q >= nb * (nb + 1) / 2) return;
    int I, J;
    lower_tile(q, I, J);
    const int a0 = 64 * I, b0 = 64 * J;
    const double *P = A.P + S.poff, *F = A.S + S.soff;
    double *U = A.U + S.uoff;
    __shared__ double As[kMk][kMpad];
    __shared__ double Bs[kMk][kMpad];
    if constexpr (MODE != 2) {
        const int nch = (w + kMk - 1) / kMk, cb = z * nch / ks, ce = (z + 1) * nch / ks;
        mma_pipeline(
            ce - cb,
            [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i, x = e % 64, k = (cb + c) * kMk + e / 64;
                    ra[i] = k < w && a0 + x < no ? P[(long long)k * m + w + a0 + x] : 0.0;
                    rb[i] = k < w && b0 + x < no ? F[(long long)k * no + b0 + x] : 0.0;
                }
            },
            [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i;
                    As[e / 64][e % 64] = ra[i];
                    Bs[e / 64][e % 64] = rb[i];
                }
            },
            As, Bs, acc);
    }
    if constexpr (MODE == 1) {
        acc_store(A.Wk + ((long long)tile * ks + z) * kAccVals, acc);
        return;
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int aa = a0 + mma_row(x, r), bb = b0 + mma_col(y);
                if (aa < no && bb < no && bb <= aa) U[(long long)bb * no + aa] -= acc.c[x][y][r];
            }
}

// Step s of the right-looking loop, first launch (`per` tiles per supernode; tiles q < split
// are row blocks, the others X blocks). Row block s + 1 + q below diagonal block s, as a product
// with the diagonal block's inverse that k_fac_potrf left in T: L(r, kb + j) = sum_t A(r, kb + t)
// X(kb + j, kb + t) (L_below = A_below L_ss^{-T}); A(i, t) from the panel (column-major), B(t, j)
// = X(kb + j, kb + t) (row-major X). X block (s, J), J = q - split < s: X_sJ = -X_ss W_sJ, where
// W_sJ = sum_{K=J}^{s-1} L_sK X_KJ, accumulated in X's place by the second launches of the
// earlier steps. Each tile is read whole before it is written.
__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_fac_trsm(FacArgs A, int h0, int s, int split, int per) {
    const FacDev S = A.sn[A.hlist[h0 + blockIdx.x / per]];
    const int w = S.w, m = S.m, kb = s * kFb, q = (int)(blockIdx.x % per);
    MD_G double *P = A.P + S.poff;
    MD_G double *T = A.T + S.toff;
    __shared__ double As[kMk][kMpad];
    __shared__ double Bs[kMk][kMpad];
    MmaAcc acc;
    mma_zero(acc);
    if (q < split) {
        const int r0 = (s + 1 + q) * kFb;
        if (r0 >= w) return;  // also kb + kFb <= r0 < w: the diagonal block is full
        mma_pipeline(
            kFb / kMk,
            [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i, x = e % 64, k0 = c * kMk;
                    ra[i] = r0 + x < w ? P[(long long)(kb + k0 + e / 64) * m + r0 + x] : 0.0;
                    // B(t, j) = X(kb + j, kb + t): consecutive threads read consecutive t of one row
                    rb[i] = T[(long long)(kb + e / kMk) * w + kb + k0 + e % kMk];
                }
            },
            [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i;
                    As[e / 64][e % 64] = ra[i];
                    Bs[e % kMk][e / kMk] = rb[i];
                }
            },
            As, Bs, acc);
        __syncthreads();  // every wave has staged its last reads of the tile before any write
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rr = r0 + mma_row(x, r), jj = kb + mma_col(y);
                    if (rr < w) P[(long long)jj * m + rr] = acc.c[x][y][r];
                }
        return;
    }
    const int J = q - split, c0 = J * kFb;
    if (J >= s || kb >= w) return;  // block J < s is full
    const int bs = min(kFb, w - kb);
    mma_pipeline(
        kFb / kMk,
        [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int e = threadIdx.x + kFacBlock * i, k0 = c * kMk;
                // A(i, t) = X(kb + i, kb + t): consecutive threads read consecutive t of one row
                const int ia = e / kMk, ta = k0 + e % kMk;
                ra[i] = ia < bs && ta < bs ? T[(long long)(kb + ia) * w + kb + ta] : 0.0;
                // B(t, j) = W(kb + t, c0 + j)
                const int tb = k0 + e / 64;
                rb[i] = tb < bs ? T[(long long)(kb + tb) * w + c0 + e % 64] : 0.0;
            }
        },
        [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int e = threadIdx.x + kFacBlock * i;
                As[e % kMk][e / kMk] = ra[i];
                Bs[e / 64][e % 64] = rb[i];
            }
        },
        As, Bs, acc);
    __syncthreads();  // W_sJ is read whole before X_sJ replaces it
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = kb + mma_row(x, r), cc = c0 + mma_col(y);
                if (rr < w) T[(long long)rr * w + cc] = -acc.c[x][y][r];
            }
}

// Step s, second launch (`per` tiles per supernode; tiles q < split are trailing tiles, the
// others W tiles). Lower trailing tile q = (I, J), s < J <= I: A(r, c) -= sum_{t<64} L(r, kb + t)
// L(c, kb + t). W tile (I, J), I = s + 1 + (q - split) / (s + 1), J = (q - split) % (s + 1):
// W_IJ += L_Is X_sJ (stored by the first contribution, J = s), for the X blocks of row block I.
// Tile 0 = (s + 1, s + 1) completes diagonal block s + 1, and its workgroup then factors and
// inverts that block (pivot_block): no other tile of the launch touches it or its block of T, so
// the next step's chain starts while the other tiles still run.
__global__ __launch_bounds__(kFacBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_fac_syrk(FacArgs A, int h0, int s, int split, int per) {
    const FacDev S = A.sn[A.hlist[h0 + blockIdx.x / per]];
    const int w = S.w, m = S.m, kb = s * kFb, q = (int)(blockIdx.x % per);
    MD_G double *P = A.P + S.poff;
    MD_G double *T = A.T + S.toff;
    __shared__ double pan[kPbPan];
    __shared__ double lsh[kPbL];
    double (*As)[kMpad] = reinterpret_cast<double (*)[kMpad]>(pan);
    double (*Bs)[kMpad] = reinterpret_cast<double (*)[kMpad]>(lsh);
    MmaAcc acc;
    mma_zero(acc);
    if (q < split) {
        const int nt = (w + kFb - 1) / kFb - 1 - s;  // row blocks below diagonal block s
        if (q >= nt * (nt + 1) / 2) return;
        int I, J;
        lower_tile(q, I, J);
        const int r0 = (s + 1 + I) * kFb, c0 = (s + 1 + J) * kFb;
        mma_pipeline(
            kFb / kMk,
            [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i, x = e % 64;
                    const long long col = (long long)(kb + c * kMk + e / 64) * m;
                    ra[i] = r0 + x < w ? P[col + r0 + x] : 0.0;
                    rb[i] = c0 + x < w ? P[col + c0 + x] : 0.0;
                }
            },
            [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
                for (int i = 0; i < kPer; ++i) {
                    const int e = threadIdx.x + kFacBlock * i;
                    As[e / 64][e % 64] = ra[i];
                    Bs[e / 64][e % 64] = rb[i];
                }
            },
            As, Bs, acc);
        if (q == 0) {
            // tile (s + 1, s + 1): its entries go to LDS for the pivot block (which writes L over
            // their place in the panel), after every wave's last read of the staged operands
            double v[2][2][4];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int rr = r0 + mma_row(x, r), cc = c0 + mma_col(y);
                        v[x][y][r] = rr < w && cc < w && rr >= cc ? P[(long long)cc * m + rr] - acc.c[x][y][r] : 0.0;
                    }
            __syncthreads();
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int r = 0; r < 4; ++r) lsh[mma_row(x, r) * kLdsPad + mma_col(y)] = v[x][y][r];
            __syncthreads();
            pivot_block<true>(A, S, s + 1, pan, lsh);
            return;
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rr = r0 + mma_row(x, r), cc = c0 + mma_col(y);
                    if (rr < w && cc < w && rr >= cc) P[(long long)cc * m + rr] -= acc.c[x][y][r];
                }
        return;
    }
    const int qq = q - split, r0 = (s + 1 + qq / (s + 1)) * kFb, J = qq % (s + 1), c0 = J * kFb;
    if (r0 >= w) return;  // also block s (and every block J <= s) is full
    mma_pipeline(
        kFb / kMk,
        [&](int c, double (&ra)[kPer], double (&rb)[kPer]) {
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int e = threadIdx.x + kFacBlock * i, x = e % 64, t = c * kMk + e / 64;
                ra[i] = r0 + x < w ? P[(long long)(kb + t) * m + r0 + x] : 0.0;  // L(r0 + x, kb + t)
                rb[i] = T[(long long)(kb + t) * w + c0 + x];                       // X(kb + t, c0 + x)
            }
        },
        [&](const double (&ra)[kPer], const double (&rb)[kPer]) {
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int e = threadIdx.x + kFacBlock * i;
                As[e / 64][e % 64] = ra[i];
                Bs[e / 64][e % 64] = rb[i];
            }
        },
        As, Bs, acc);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = r0 + mma_row(x, r), cc = c0 + mma_col(y);
                if (rr < w) {
                    MD_G double *dst = T + (long long)rr * w + cc;
                    *dst = (J == s ? 0.0 : *dst) + acc.c[x][y][r];
                }
            }
}

// A whole height of small fronts (every m = w + noff <= mm <= kSmallFront) in one launch, one
// wavefront per front, the front in LDS (2 mm^2 doubles, sized by the launch): assembly (the
// original entries, then the children's update matrices in child order, as k_fac_assemble adds
// them; the front's row list staged in LDS for the searches), the pivot block's Cholesky
// (right-looking), X = L^{-1} by columns (one lane per column), D^{-1} = X^T X, G = F_os D^{-1},
// U = U - G F_os^T into the update workspace for the parent, and the two panel layouts in the
// solve precision. The FP64 panel workspace is not used.
constexpr int kSmallFront = 64;
constexpr int kSmallBlock = 64;
constexpr int kSmallSlotLanes = 4;  // lanes per pivot column for its original entries
__global__ __launch_bounds__(kSmallBlock) void k_fac_small(FacArgs A, int h0, int mm, int mw) {
    const FacDev S = A.sn[A.hlist[h0 + blockIdx.x]];
    const int tid = threadIdx.x, w = S.w, no = S.noff, m = S.m;
    if (m > mm || w > mw || mm > kSmallFront) {  // the host sends only heights of small fronts
        if (tid == 0) *A.flag = 2;
        return;
    }
    extern __shared__ double smem[];
    double *Fs = smem;            // the front (lower, row stride mm), then L / D^{-1}
    double *Xs = smem + mm * mm;  // m x w (row stride mw): X (rows < w), then G (rows >= w)
    __shared__ int rows[kSmallFront];
    for (int e = tid; e < no; e += kSmallBlock) rows[e] = A.offrows[S.roff + e];
    for (int e = tid; e < m * mm; e += kSmallBlock) Fs[e] = 0.0;
    __syncthreads();
    // original entries, kSmallSlotLanes lanes per pivot column (one slot each; every front entry
    // receives at most one)
    for (int k = tid / kSmallSlotLanes; k < w; k += kSmallBlock / kSmallSlotLanes) {
        const int j = S.c0 + k, sl = j / kSlice, l = j % kSlice;
        if (tid % kSmallSlotLanes == 0) Fs[k * mm + k] = A.diag[j];
        for (int q = A.slice_ptr[sl] + tid % kSmallSlotLanes; q < A.slice_ptr[sl + 1]; q += kSmallSlotLanes) {
            const long long g = (long long)q * kSlice + l;
            const int i = A.sell_col[g];
            if (i <= j || i >= A.nf) continue;
            const double v = -A.w64[g];
            if (i < S.c0 + w) {
                Fs[(i - S.c0) * mm + k] += v;
            } else {
                int lo = 0, hi = no;  // first off row >= i
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (rows[mid] < i) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < no && rows[lo] == i) Fs[(w + lo) * mm + k] += v;
                else *A.flag = 2;
            }
        }
    }
    __syncthreads();
    // children's update matrices in child order (distinct entries of one child land on distinct
    // front entries)
    for (int ci = S.cb; ci < S.ce; ++ci) {
        const FacDev C = A.sn[A.child[ci]];
        const int nc = C.noff;
        const MD_G int *rel = A.rel + C.roff;
        const MD_G double *Uc = A.U + C.uoff;
        for (int e = tid; e < nc * nc; e += kSmallBlock) {
            const int b = e / nc, a = e - b * nc;
            if (a < b) continue;
            const int ra = rel[a], rb = rel[b];
            if (ra >= m || rb > ra) { *A.flag = 2; continue; }
            Fs[ra * mm + rb] += Uc[(long long)b * nc + a];
        }
        __syncthreads();
    }
    // Cholesky of the pivot block (1 / sqrt(pivot): v_rsq_f64 and two Newton steps, kept for X)
    const double amax = __longlong_as_double((long long)*A.amax);
    __shared__ double rdg[kSmallFront];
    for (int k = 0; k < w; ++k) {
        const double piv = Fs[k * mm + k];
        if (tid == 0 && !(piv > 1e-14 * amax)) *A.flag = 1;
        double rd = __builtin_amdgcn_rsq(piv);
        rd = rd * (1.5 - 0.5 * piv * rd * rd);
        rd = rd * (1.5 - 0.5 * piv * rd * rd);
        const double d = piv * rd;
        __syncthreads();  // every lane has read the pivot before it changes
        if (tid == 0) {
            Fs[k * mm + k] = d;
            rdg[k] = rd;
        }
        for (int i = k + 1 + tid; i < w; i += kSmallBlock) Fs[i * mm + k] *= rd;
        __syncthreads();
        const int t = w - k - 1;  // trailing lower triangle, t x t
        for (int e = tid; e < t * t; e += kSmallBlock) {
            const int i = k + 1 + e / t, j = k + 1 + e % t;
            if (j <= i) Fs[i * mm + j] -= Fs[i * mm + k] * Fs[j * mm + k];
        }
        __syncthreads();
    }
    // X = L^{-1}, column i by lane i: X(r, i) = (delta_ri - sum_{i<=t<r} L(r, t) X(t, i)) / L(r, r)
    for (int i = tid; i < w; i += kSmallBlock) {
        for (int r = 0; r < i; ++r) Xs[r * mw + i] = 0.0;
        for (int r = i; r < w; ++r) {
            double acc = r == i ? 1.0 : 0.0;
            for (int t = i; t < r; ++t) acc -= Fs[r * mm + t] * Xs[t * mw + i];
            Xs[r * mw + i] = acc * rdg[r];
        }
    }
    __syncthreads();
    // D^{-1}(r, c) = sum_{k >= max(r, c)} X(k, r) X(k, c), over the pivot block (L is done)
    for (int e = tid; e < w * w; e += kSmallBlock) {
        const int r = e / w, c = e % w;
        double acc = 0.0;
        for (int k = max(r, c); k < w; ++k) acc += Xs[k * mw + r] * Xs[k * mw + c];
        Fs[r * mm + c] = acc;  // the upper part too: D^{-1} is symmetric
    }
    __syncthreads();
    // G(r, c) = sum_k F_os(r, k) D^{-1}(k, c) into Xs rows w + r (X is no longer needed)
    for (int e = tid; e < no * w; e += kSmallBlock) {
        const int r = e / w, c = e % w;
        double acc = 0.0;
        for (int k = 0; k < w; ++k) acc += Fs[(w + r) * mm + k] * Fs[k * mm + c];
        Xs[(w + r) * mw + c] = acc;
    }
    __syncthreads();
    // U(a, b) = U_front(a, b) - sum_k G(a, k) F_os(b, k), lower part, column-major for the parent
    MD_G double *U = A.U + S.uoff;
    for (int e = tid; e < no * no; e += kSmallBlock) {
        const int b = e / no, a = e - b * no;
        if (a < b) continue;
        double acc = Fs[(w + a) * mm + w + b];
        for (int k = 0; k < w; ++k) acc -= Xs[(w + a) * mw + k] * Fs[(w + b) * mm + k];
        U[(long long)b * no + a] = acc;
    }
    // the sweeps' panels: row-major [D^{-1}; G] (m x w) and column-major G (no x w)
    for (int e = tid; e < m * w; e += kSmallBlock) {
        const int r = e / w, c = e % w;
        const double v = r < w ? Fs[r * mm + c] : Xs[r * mw + c];
        if (A.solve64) static_cast<MD_G double *>(A.pr)[S.pr_off + e] = v;
        else static_cast<MD_G float *>(A.pr)[S.pr_off + e] = (float)v;
    }
    for (int e = tid; e < no * w; e += kSmallBlock) {
        const int c = e / no, r = e - c * no;
        const double v = Xs[(w + r) * mw + c];
        if (A.solve64) static_cast<MD_G double *>(A.pc)[S.pc_off + e] = v;
        else static_cast<MD_G float *>(A.pc)[S.pc_off + e] = (float)v;
    }
}

// Plain device copy of n 16-byte elements (grid-stride): the known-size reference kernel for
// HBM counters and the achievable copy bandwidth.
__global__ __launch_bounds__(kBlock) void k_copy_probe(const f4 *__restrict__ src, f4 *__restrict__ dst, long long n) {
    const long long stride = (long long)gridDim.x * kBlock;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

void launch_fac_diag(hipStream_t s, const FacArgs &A) {
    if (A.nf > 0) hipLaunchKernelGGL(k_fac_diag, dim3((A.nf + kBlock - 1) / kBlock), dim3(kBlock), 0, s, A);
}
void launch_fac_small(hipStream_t s, const FacArgs &A, int h0, int count, int mm, int mw) {
    if (count > 0)
        hipLaunchKernelGGL(k_fac_small, dim3(count), dim3(kSmallBlock), sizeof(double) * (mm * mm + mm * mw), s, A, h0,
                           mm, mw);
}
void launch_fac_height(hipStream_t s, const FacArgs &A, int h0, int count, int nb, int asm_cols, int per_asm,
                       const FacSplit &dinv, const FacSplit &g, const FacSplit &u) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fac_assemble, dim3(count * per_asm), dim3(kFacBlock), 0, s, A, h0, per_asm, asm_cols);
    hipLaunchKernelGGL(k_fac_potrf, dim3(count), dim3(kFacBlock), 0, s, A, h0, 0);
    // per step b: the row blocks and X blocks (X_bJ, J < b) of diagonal block b, then the trailing
    // tiles and the W tiles of the row blocks below (W_IJ, I > b, J <= b), with diagonal block b + 1
    for (int b = 0; b < nb; ++b) {
        // diagonal block b is factored: by k_fac_potrf (b = 0) or by tile 0 of step b - 1's k_fac_syrk
        const int below = nb - 1 - b;
        if (below + b > 0)
            hipLaunchKernelGGL(k_fac_trsm, dim3(count * (below + b)), dim3(kFacBlock), 0, s, A, h0, b, below, below + b);
        const int tiles = below * (below + 1) / 2, wtiles = below * (b + 1);
        if (tiles + wtiles > 0)
            hipLaunchKernelGGL(k_fac_syrk, dim3(count * (tiles + wtiles)), dim3(kFacBlock), 0, s, A, h0, b, tiles,
                               tiles + wtiles);
    }
    // the products, split over K where ks > 1
    auto product = [&](auto plain, auto part, auto fix, const FacSplit &p) {
        if (p.ks > 1) {
            hipLaunchKernelGGL(part, dim3(count * p.per * p.ks), dim3(kFacBlock), 0, s, A, h0, p.per, p.ks);
            hipLaunchKernelGGL(fix, dim3(count * p.per), dim3(kFacBlock), 0, s, A, h0, p.per, p.ks);
        } else {
            hipLaunchKernelGGL(plain, dim3(count * p.per), dim3(kFacBlock), 0, s, A, h0, p.per, 1);
        }
    };
    product(k_fac_dinv<0>, k_fac_dinv<1>, k_fac_dinv<2>, dinv);
    if (g.per > 0) {
        product(k_fac_g<0>, k_fac_g<1>, k_fac_g<2>, g);
        product(k_fac_u<0>, k_fac_u<1>, k_fac_u<2>, u);
    }
}

void launch_copy_probe(hipStream_t s, const void *src, void *dst, long long n16) {
    if (n16 <= 0) return;
    hipLaunchKernelGGL(k_copy_probe, dim3(4096), dim3(kBlock), 0, s, static_cast<const f4 *>(src), static_cast<f4 *>(dst), n16);
}

#define MD_LAUNCH_T(kern, blocks, s, ...)                                                        \
    do {                                                                                        \
        if ((blocks) <= 0) return;                                                              \
        if (fp64) hipLaunchKernelGGL(kern<double>, dim3(blocks), dim3(kBlock), 0, s, __VA_ARGS__); \
        else hipLaunchKernelGGL(kern<float>, dim3(blocks), dim3(kBlock), 0, s, __VA_ARGS__);     \
    } while (0)

void launch_pcg_diag(hipStream_t s, LaunchSet L, int blocks, bool fp64) { MD_LAUNCH_T(k_pcg_diag, blocks, s, L); }
void launch_pcg_init(hipStream_t s, LaunchSet L, int blocks, bool fp64) { MD_LAUNCH_T(k_pcg_init, blocks, s, L); }
void launch_pcg_step(hipStream_t s, LaunchSet L, int blocks, int step, bool fp64) {
    MD_LAUNCH_T(k_pcg_step, blocks, s, L, step);
}
void launch_pcg_apply(hipStream_t s, LaunchSet L, int blocks, bool fp64) { MD_LAUNCH_T(k_pcg_apply, blocks, s, L); }

void launch_normals(hipStream_t s, LaunchSet Lf, int blocks_face, LaunchSet Ls, int blocks_slice, double *fn, d4 *out) {
    if (blocks_face > 0) hipLaunchKernelGGL(k_face_normals, dim3(blocks_face), dim3(kBlock), 0, s, Lf, fn);
    if (blocks_slice > 0) hipLaunchKernelGGL(k_vertex_normals, dim3(blocks_slice), dim3(kBlock), 0, s, Ls, (const double *)fn, out);
}

hipError_t scratch_address(void **ptr) {
    *ptr = nullptr;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

void launch_pack_xyz(hipStream_t s, const d4 *v, const int *perm, int n, double *dst) {
    if (n > 0) hipLaunchKernelGGL(k_pack_xyz, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, v, perm, n, dst);
}
void launch_unpack_xyz(hipStream_t s, const double *xyz, const int *perm, int n, int nf, const double *tgt, d4 *p,
                       d4 *pp) {
    if (n > 0)
        hipLaunchKernelGGL(k_unpack_xyz, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, xyz, perm, n, nf, tgt, p,
                           pp);
}

void launch_scatter_entries(hipStream_t s, const TargetEntry *e, int k) {
    if (k > 0) hipLaunchKernelGGL(k_scatter_entries, dim3((k + kBlock - 1) / kBlock), dim3(kBlock), 0, s, e, k);
}

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
    if (blocks <= kPreloadMaxTiles) {
        if (fp64) hipLaunchKernelGGL((k_fwd<double, true>), dim3(blocks), dim3(kTile), 0, s, L, level);
        else hipLaunchKernelGGL((k_fwd<float, true>), dim3(blocks), dim3(kTile), 0, s, L, level);
    } else {
        if (fp64) hipLaunchKernelGGL((k_fwd<double, false>), dim3(blocks), dim3(kTile), 0, s, L, level);
        else hipLaunchKernelGGL((k_fwd<float, false>), dim3(blocks), dim3(kTile), 0, s, L, level);
    }
}
void launch_bwd(hipStream_t s, LaunchSet L, int blocks, int level, bool fp64) {
    if (blocks <= 0) return;
    if (blocks <= kPreloadMaxTiles) {
        if (fp64) hipLaunchKernelGGL((k_bwd<double, true>), dim3(blocks), dim3(kTile), 0, s, L, level);
        else hipLaunchKernelGGL((k_bwd<float, true>), dim3(blocks), dim3(kTile), 0, s, L, level);
    } else {
        if (fp64) hipLaunchKernelGGL((k_bwd<double, false>), dim3(blocks), dim3(kTile), 0, s, L, level);
        else hipLaunchKernelGGL((k_bwd<float, false>), dim3(blocks), dim3(kTile), 0, s, L, level);
    }
}
void launch_sub_fwd(hipStream_t s, LaunchSet L, int blocks, int nodes, int lds_entries, bool fused, bool fp64) {
    if (blocks <= 0) return;
    const size_t rec = sizeof(SubNode) * (size_t)nodes;
    const size_t lds = rec + (fp64 ? sizeof(d4) : sizeof(f4)) * lds_entries;
    if (fused) {
        if (fp64) hipLaunchKernelGGL((k_sub_fwd<double, true>), dim3(blocks), dim3(kTile), lds, s, L, nodes);
        else hipLaunchKernelGGL((k_sub_fwd<float, true>), dim3(blocks), dim3(kTile), lds, s, L, nodes);
    } else {
        if (fp64) hipLaunchKernelGGL((k_sub_fwd<double, false>), dim3(blocks), dim3(kTile), lds, s, L, nodes);
        else hipLaunchKernelGGL((k_sub_fwd<float, false>), dim3(blocks), dim3(kTile), lds, s, L, nodes);
    }
}
void launch_sub_bwd(hipStream_t s, LaunchSet L, int blocks, int nodes, int lds_entries, bool fp64) {
    if (blocks <= 0) return;
    const size_t rec = sizeof(SubNode) * (size_t)nodes;
    if (fp64) hipLaunchKernelGGL(k_sub_bwd<double>, dim3(blocks), dim3(kTile), rec + sizeof(d4) * lds_entries, s, L, nodes);
    else hipLaunchKernelGGL(k_sub_bwd<float>, dim3(blocks), dim3(kTile), rec + sizeof(f4) * lds_entries, s, L, nodes);
}
void launch_energy(hipStream_t s, LaunchSet L, int blocks, bool fp64) {
    if (blocks <= 0) return;
    if (fp64) hipLaunchKernelGGL(k_energy<double>, dim3(blocks), dim3(kBlock), 0, s, L);
    else hipLaunchKernelGGL(k_energy<float>, dim3(blocks), dim3(kBlock), 0, s, L);
}

}  // namespace md

namespace md {
void set_error(const std::string &msg);
}

// Host diagnostic: the local step's rotation code compiled for the host (correctly rounded
// reciprocal and square roots instead of the device's approximate ones).
extern "C" int md_host_best_rotation(int fp64, int count, const double *cov, const double *prev, double *quat,
                                     int *fallback) {
    if (count < 0 || (count > 0 && (!cov || !quat))) {
        md::set_error("md_host_best_rotation: invalid argument");
        return MD_ERR_INVALID;
    }
    for (int k = 0; k < count; ++k) {
        const double *c = cov + 9 * (size_t)k;
        const double p[4] = {prev ? prev[4 * k] : 0.0, prev ? prev[4 * k + 1] : 0.0, prev ? prev[4 * k + 2] : 0.0,
                             prev ? prev[4 * k + 3] : 1.0};
        bool qcp;
        if (fp64) {
            double m[9], q[4];
            for (int i = 0; i < 9; ++i) m[i] = c[i];
            qcp = md::best_rotation<double>(m, p[0], p[1], p[2], p[3], q[0], q[1], q[2], q[3]);
            for (int i = 0; i < 4; ++i) quat[4 * k + i] = q[i];
        } else {
            float m[9], q[4];
            for (int i = 0; i < 9; ++i) m[i] = (float)c[i];
            qcp = md::best_rotation<float>(m, (float)p[0], (float)p[1], (float)p[2], (float)p[3], q[0], q[1], q[2], q[3]);
            for (int i = 0; i < 4; ++i) quat[4 * k + i] = q[i];
        }
        if (fallback) fallback[k] = qcp ? 0 : 1;
    }
    return MD_OK;
}
