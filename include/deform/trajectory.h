/*
 * deform/trajectory.h -- 6-DoF keyframe trajectory, counterpart of inc/deform/trajectory.h
 * of cheind/mesh-deform (TrajectorySE3, trajectory.h:33-83), computed without Eigen or Sophus.
 *
 * Evaluation follows the reference: each keyframe is mapped to se(3) by the SE3 logarithm
 * (tangent (upsilon, omega), Sophus' convention), a cubic B-spline interpolates the tangents at
 * normalised chord-length parameters with averaged knots (Eigen's
 * SplineFitting<..>::Interpolate(points, 3)), and the spline value is mapped back by the SE3
 * exponential. With four keyframes the spline is one cubic Bezier segment.
 *
 * Transform types. The reference takes and returns Eigen::Transform<Scalar, 3, Affine>
 * (trajectory.h:46,55-73). Here the transform type is a template parameter: anything that
 * reads and writes as a 4x4 affine matrix -- t.matrix()(r, c) (Eigen::Transform and Eigen
 * matrices), t.m[r][c] (deform::Transform below) or t(r, c) -- and has a static Identity().
 * The default is Eigen::Transform<Scalar, 3, Eigen::Affine> where Eigen is on the include path,
 * so code written against the reference compiles unchanged, and deform::Transform<Scalar>
 * otherwise. Only the affine rows 0-2 are read and written; the transforms are rigid.
 */
#ifndef DEFORM_TRAJECTORY_H
#define DEFORM_TRAJECTORY_H

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#if __has_include(<Eigen/Geometry>)
#include <Eigen/Geometry>
#define DEFORM_HAVE_EIGEN_TRANSFORM 1
#endif

namespace deform {

/** Row-major 4x4 rigid transform. */
template <class S> struct Transform {
    S m[4][4];
    static Transform Identity() {
        Transform t;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) t.m[r][c] = r == c ? S(1) : S(0);
        return t;
    }
    Transform operator*(const Transform &o) const {
        Transform t;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                S a = 0;
                for (int k = 0; k < 4; ++k) a += m[r][k] * o.m[k][c];
                t.m[r][c] = a;
            }
        return t;
    }
    /** Inverse of a rigid transform (the Isometry inverse of deformation_util.h:38). */
    Transform isometryInverse() const {
        Transform t = Identity();
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) t.m[r][c] = m[c][r];
        for (int r = 0; r < 3; ++r) t.m[r][3] = -(t.m[r][0] * m[0][3] + t.m[r][1] * m[1][3] + t.m[r][2] * m[2][3]);
        return t;
    }
    std::array<S, 3> apply(const std::array<S, 3> &p) const {
        std::array<S, 3> q;
        for (int r = 0; r < 3; ++r) q[r] = m[r][0] * p[0] + m[r][1] * p[1] + m[r][2] * p[2] + m[r][3];
        return q;
    }
};

namespace detail {

// Priority tags: the first matching access form wins.
struct xf_p0 {};
struct xf_p1 : xf_p0 {};
struct xf_p2 : xf_p1 {};

template <class T> auto xf_at(T &t, int r, int c, xf_p2) -> decltype(t.matrix()(r, c)) { return t.matrix()(r, c); }
template <class T> auto xf_at(T &t, int r, int c, xf_p1) -> decltype(t.m[r][c]) { return t.m[r][c]; }
template <class T> auto xf_at(T &t, int r, int c, xf_p0) -> decltype(t(r, c)) { return t(r, c); }

/** Entry (r, c) of a 4x4 affine transform of the caller's type (a reference when t is mutable). */
template <class T> auto xf_at(T &t, int r, int c) -> decltype(xf_at(t, r, c, xf_p2())) {
    return xf_at(t, r, c, xf_p2());
}

/** The caller's transform as a deform::Transform (affine rows 0-2; row 3 is [0 0 0 1]). */
template <class S, class T> Transform<S> to_rigid(const T &t) {
    Transform<S> o = Transform<S>::Identity();
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) o.m[r][c] = (S)xf_at(t, r, c);
    return o;
}

/** A deform::Transform as the caller's transform type (T::Identity() with rows 0-2 set). */
template <class T, class S> T from_rigid(const Transform<S> &x) {
    T t = T::Identity();
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) {
            auto &e = xf_at(t, r, c);
            e = (typename std::decay<decltype(e)>::type)x.m[r][c];
        }
    return t;
}

}  // namespace detail

#ifdef DEFORM_HAVE_EIGEN_TRANSFORM
template <class S> using DefaultTransform = Eigen::Transform<S, 3, Eigen::Affine>;
#else
template <class S> using DefaultTransform = Transform<S>;
#endif

namespace se3 {

template <class S> void hat(const S w[3], S W[3][3]) {
    W[0][0] = 0; W[0][1] = -w[2]; W[0][2] = w[1];
    W[1][0] = w[2]; W[1][1] = 0; W[1][2] = -w[0];
    W[2][0] = -w[1]; W[2][1] = w[0]; W[2][2] = 0;
}

template <class S> void mat3_add_scaled(S out[3][3], const S A[3][3], S a, const S B[3][3], S b) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[r][c] += a * A[r][c] + b * B[r][c];
}

template <class S> void sq(const S W[3][3], S W2[3][3]) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) W2[r][c] = W[r][0] * W[0][c] + W[r][1] * W[1][c] + W[r][2] * W[2][c];
}

/** exp: tangent (upsilon, omega) -> transform. */
template <class S> Transform<S> exp(const S xi[6]) {
    const S *ups = xi, *w = xi + 3;
    double th = std::sqrt((double)w[0] * w[0] + (double)w[1] * w[1] + (double)w[2] * w[2]);
    S W[3][3], W2[3][3];
    hat(w, W);
    sq(W, W2);
    S a, b, c, d;  // R = I + a W + b W^2, V = I + c W + d W^2
    if (th < 1e-10) { a = 1; b = 0.5; c = 0.5; d = S(1.0 / 6.0); }
    else {
        a = S(std::sin(th) / th);
        b = S((1 - std::cos(th)) / (th * th));
        c = b;
        d = S((th - std::sin(th)) / (th * th * th));
    }
    Transform<S> T = Transform<S>::Identity();
    S R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}, V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    mat3_add_scaled(R, W, a, W2, b);
    mat3_add_scaled(V, W, c, W2, d);
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) T.m[r][k] = R[r][k];
        T.m[r][3] = V[r][0] * ups[0] + V[r][1] * ups[1] + V[r][2] * ups[2];
    }
    return T;
}

/** log: transform -> tangent (upsilon, omega), rotation angle in [0, pi]. */
template <class S> void log(const Transform<S> &T, S xi[6]) {
    double tr = (double)T.m[0][0] + T.m[1][1] + T.m[2][2];
    double cth = std::max(-1.0, std::min(1.0, (tr - 1.0) * 0.5));
    double th = std::acos(cth);
    double v[3] = {(double)T.m[2][1] - T.m[1][2], (double)T.m[0][2] - T.m[2][0], (double)T.m[1][0] - T.m[0][1]};
    S w[3];
    if (th < 1e-7) {
        for (int k = 0; k < 3; ++k) w[k] = S(0.5 * v[k]);
    } else if (M_PI - th < 1e-6) {
        int k = 0;
        for (int j = 1; j < 3; ++j)
            if (T.m[j][j] > T.m[k][k]) k = j;
        double akk = (T.m[k][k] + 1.0) * 0.5;
        for (int j = 0; j < 3; ++j) w[j] = S(((T.m[j][k] + (j == k ? 1.0 : 0.0)) * 0.5) / std::sqrt(std::max(akk, 1e-300)) * th);
    } else {
        for (int k = 0; k < 3; ++k) w[k] = S(th / (2.0 * std::sin(th)) * v[k]);
    }
    S W[3][3], W2[3][3];
    hat(w, W);
    sq(W, W2);
    S b = th < 1e-10 ? S(1.0 / 12.0) : S((1 - th / (2 * std::tan(th / 2))) / (th * th));
    S Vi[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    mat3_add_scaled(Vi, W, S(-0.5), W2, b);
    for (int r = 0; r < 3; ++r) xi[r] = Vi[r][0] * T.m[0][3] + Vi[r][1] * T.m[1][3] + Vi[r][2] * T.m[2][3];
    for (int k = 0; k < 3; ++k) xi[3 + k] = w[k];
}

}  // namespace se3

namespace spline {

/** Values of the n = knots.size() - p - 1 B-spline basis functions of degree p at u. */
inline std::vector<double> basis(const std::vector<double> &knots, int p, double u) {
    int nk = (int)knots.size();
    std::vector<double> N(nk - 1, 0.0);
    if (u >= knots.back()) {
        int span = 0;
        for (int i = 0; i < nk - 1; ++i)
            if (knots[i] < knots.back()) span = i;
        N[span] = 1.0;
    } else {
        for (int i = 0; i < nk - 1; ++i)
            if (knots[i] <= u && u < knots[i + 1]) N[i] = 1.0;
    }
    for (int d = 1; d <= p; ++d) {
        std::vector<double> M(nk - 1 - d, 0.0);
        for (int i = 0; i < (int)M.size(); ++i) {
            double a = 0.0, den = knots[i + d] - knots[i];
            if (den > 0) a += (u - knots[i]) / den * N[i];
            den = knots[i + d + 1] - knots[i + 1];
            if (den > 0) a += (knots[i + d + 1] - u) / den * N[i + 1];
            M[i] = a;
        }
        N.swap(M);
    }
    N.resize(nk - p - 1);
    return N;
}

}  // namespace spline

/** Keyframe trajectory interpolated in se(3) (trajectory.h:33-83); key poses and results are of
 *  the caller's transform type (Eigen::Transform<Scalar, 3, Affine> by default where Eigen is
 *  available, as in the reference). */
template <class PrecisionType, class TransformType = DefaultTransform<PrecisionType>> class TrajectorySE3 {
public:
    typedef PrecisionType Scalar;
    typedef TransformType Transform;

    TrajectorySE3() : _dirty(false), _p(3) {}

    /** addKeyPose (trajectory.h:55-59): appends a key pose and returns it. */
    Transform addKeyPose(const Transform &t) {
        _poses.push_back(detail::to_rigid<Scalar>(t));
        _dirty = true;
        return t;
    }

    /** operator()(time) (trajectory.h:61-73): the pose at time in [0, 1]. */
    Transform operator()(Scalar time) { return detail::from_rigid<Transform>(rigid(time)); }

    /** The pose at time as a deform::Transform. */
    deform::Transform<Scalar> rigid(Scalar time) {
        if (_dirty) fit();
        std::vector<double> N = spline::basis(_knots, _p, (double)time);
        Scalar xi[6];
        for (int d = 0; d < 6; ++d) {
            double a = 0.0;
            for (size_t j = 0; j < N.size(); ++j) a += N[j] * _ctrl[j][d];
            xi[d] = (Scalar)a;
        }
        return se3::exp(xi);
    }

private:
    void fit() {
        int n = (int)_poses.size();
        if (n < 2) throw std::runtime_error("TrajectorySE3: at least two key poses are needed");
        std::vector<std::array<double, 6>> pts(n);
        for (int i = 0; i < n; ++i) {
            Scalar xi[6];
            se3::log(_poses[i], xi);
            for (int d = 0; d < 6; ++d) pts[i][d] = xi[d];
        }
        _p = std::min(3, n - 1);
        // chord-length parameters (SplineFitting::ChordLengths)
        std::vector<double> u(n, 0.0);
        for (int i = 1; i < n; ++i) {
            double s = 0.0;
            for (int d = 0; d < 6; ++d) s += (pts[i][d] - pts[i - 1][d]) * (pts[i][d] - pts[i - 1][d]);
            u[i] = u[i - 1] + std::sqrt(s);
        }
        for (int i = 1; i < n; ++i) u[i] = u[n - 1] > 0 ? u[i] / u[n - 1] : double(i) / (n - 1);
        u[n - 1] = 1.0;
        // averaged knots (KnotAveraging)
        _knots.assign(n + _p + 1, 0.0);
        for (int i = n; i < n + _p + 1; ++i) _knots[i] = 1.0;
        for (int j = 1; j < n - _p; ++j) {
            double s = 0.0;
            for (int k = j; k < j + _p; ++k) s += u[k];
            _knots[j + _p] = s / _p;
        }
        // collocation system sum_j N_j(u_i) P_j = pts_i, Gaussian elimination with pivoting
        std::vector<std::vector<double>> A(n, std::vector<double>(n + 6, 0.0));
        for (int i = 0; i < n; ++i) {
            std::vector<double> N = spline::basis(_knots, _p, u[i]);
            for (int j = 0; j < n; ++j) A[i][j] = N[j];
            for (int d = 0; d < 6; ++d) A[i][n + d] = pts[i][d];
        }
        for (int c = 0; c < n; ++c) {
            int piv = c;
            for (int r = c + 1; r < n; ++r)
                if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
            std::swap(A[c], A[piv]);
            for (int r = 0; r < n; ++r) {
                if (r == c || A[r][c] == 0.0) continue;
                double f = A[r][c] / A[c][c];
                for (int k = c; k < n + 6; ++k) A[r][k] -= f * A[c][k];
            }
        }
        _ctrl.assign(n, {});
        for (int j = 0; j < n; ++j)
            for (int d = 0; d < 6; ++d) _ctrl[j][d] = A[j][n + d] / A[j][j];
        _dirty = false;
    }

    std::vector<deform::Transform<Scalar>> _poses;
    bool _dirty;
    int _p;
    std::vector<double> _knots;
    std::vector<std::array<double, 6>> _ctrl;
};

}  // namespace deform

#endif
