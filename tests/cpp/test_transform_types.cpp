// The trajectory and handle-motion helpers with a transform type that reads like Eigen's
// (t.matrix()(r, c), static Identity(), composition by operator*): the calls of the
// reference's tests/test_trajectory.cpp:17-38 and examples/deform_trajectory.cpp:53-56,84,
// written against include/deform/trajectory.h and deformation_util.h, compile unchanged
// apart from the type names, and give the expected poses and handle targets.
#include <deform/array_adapter.h>
#include <deform/deformation_util.h>
#include <deform/trajectory.h>

#include <cmath>
#include <cstdio>
#include <vector>

namespace eigenlike {

// 4x4 matrix with Eigen's element access.
template <class S> struct Matrix4 {
    S a[16];
    S &operator()(int r, int c) { return a[4 * r + c]; }
    S operator()(int r, int c) const { return a[4 * r + c]; }
};

// Affine transform with Eigen::Transform's matrix() access and composition.
template <class S> struct Affine {
    Matrix4<S> M;
    Matrix4<S> &matrix() { return M; }
    const Matrix4<S> &matrix() const { return M; }
    static Affine Identity() {
        Affine t;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) t.M(r, c) = r == c ? S(1) : S(0);
        return t;
    }
    Affine operator*(const Affine &o) const {
        Affine t;
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) {
                S v = 0;
                for (int k = 0; k < 4; ++k) v += M(r, k) * o.M(k, c);
                t.M(r, c) = v;
            }
        return t;
    }
};

template <class S> Affine<S> Translation(S x, S y, S z) {
    Affine<S> t = Affine<S>::Identity();
    t.M(0, 3) = x; t.M(1, 3) = y; t.M(2, 3) = z;
    return t;
}

template <class S> Affine<S> AngleAxisUnitX(S a) {
    Affine<S> t = Affine<S>::Identity();
    t.M(1, 1) = std::cos(a); t.M(1, 2) = -std::sin(a);
    t.M(2, 1) = std::sin(a); t.M(2, 2) = std::cos(a);
    return t;
}

// Eigen's isApprox: ||a - b|| <= prec * min(||a||, ||b||).
template <class S> bool isApprox(const Matrix4<S> &a, const Matrix4<S> &b, S prec) {
    double d = 0, na = 0, nb = 0;
    for (int k = 0; k < 16; ++k) {
        d += double(a.a[k] - b.a[k]) * (a.a[k] - b.a[k]);
        na += double(a.a[k]) * a.a[k];
        nb += double(b.a[k]) * b.a[k];
    }
    return std::sqrt(d) <= prec * std::sqrt(std::fmin(na, nb));
}

}  // namespace eigenlike

// Records what DeformationUtil asks of the ARAP object.
struct Recorder {
    std::vector<int> idx;
    std::vector<double> xyz;
    template <class V> void setConstraint(int i, const V &p) {
        idx.push_back(i);
        for (int k = 0; k < 3; ++k) xyz.push_back((double)p[k]);
    }
};

static int failures = 0;
#define REQUIRE(c)                                                    \
    do {                                                              \
        if (!(c)) {                                                   \
            std::printf("FAILED line %d: %s\n", __LINE__, #c);        \
            ++failures;                                               \
        }                                                             \
    } while (0)

int main() {
    using eigenlike::AngleAxisUnitX;
    using eigenlike::Translation;
    // tests/test_trajectory.cpp:17-38
    {
        typedef deform::TrajectorySE3<float, eigenlike::Affine<float>> Trajectory;
        Trajectory path;
        Trajectory::Transform pose0 = Trajectory::Transform::Identity();
        path.addKeyPose(pose0);
        Trajectory::Transform pose1 = Translation<float>(0, 0, 1) * pose0;
        path.addKeyPose(pose1);
        Trajectory::Transform pose2 = Translation<float>(0, 0, 1) * AngleAxisUnitX<float>((float)M_PI / 4.f) * pose1;
        path.addKeyPose(pose2);
        Trajectory::Transform pose3 = Translation<float>(0, 0, 1) * pose2;
        path.addKeyPose(pose3);
        REQUIRE(eigenlike::isApprox(path(0.f).matrix(), pose0.matrix(), 1e-3f));
        REQUIRE(eigenlike::isApprox(path(1.f).matrix(), pose3.matrix(), 1e-3f));
        // the same path with the library's own transform type gives the same poses
        deform::TrajectorySE3<float, deform::Transform<float>> own;
        for (const auto &p : {pose0, pose1, pose2, pose3}) own.addKeyPose(deform::detail::to_rigid<float>(p));
        for (float t : {0.f, 0.3f, 0.7f, 1.f}) {
            auto a = path(t);
            auto b = own(t);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) REQUIRE(a.matrix()(r, c) == b.m[r][c]);
        }
    }
    // examples/deform_trajectory.cpp:53-56,84: a float trajectory drives the handles of a
    // double mesh through DeformationUtil<Adapter>
    {
        typedef deform::TrajectorySE3<float, eigenlike::Affine<float>> Trajectory;
        typedef deform::ArrayAdapter<double> Adapter;
        typedef deform::DeformationUtil<Adapter> DeformationUtil;
        std::vector<double> xyz = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0.5, 2, 0.5, 1};
        std::vector<int> tri = {0, 1, 2, 1, 3, 2, 1, 4, 3};
        Adapter ma(xyz, tri);
        const std::vector<int> handles = {3, 4};
        Trajectory trajectory;
        Trajectory::Transform prev;
        prev = trajectory.addKeyPose(Trajectory::Transform::Identity());
        prev = trajectory.addKeyPose(prev * Translation<float>(1.f, 0.f, 0.f));
        prev = trajectory.addKeyPose(prev * Translation<float>(2.f, 0.f, 0.f));
        prev = trajectory.addKeyPose(prev * AngleAxisUnitX<float>((float)M_PI / 2.f));
        DeformationUtil dutil(ma, handles.begin(), handles.end(), trajectory(0.0));
        for (double t : {0.0, 0.25, 0.5, 1.0}) {
            Recorder arap;
            auto T = trajectory((float)t);
            dutil.updateConstraints(T, arap);
            REQUIRE(arap.idx == handles);
            // expected: O T O^-1 p0 with O = trajectory(0), in double
            auto O = trajectory(0.f);
            double Od[4][4], Td[4][4], Oi[4][4] = {{0}};
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) { Od[r][c] = O.matrix()(r, c); Td[r][c] = T.matrix()(r, c); }
            for (int r = 0; r < 3; ++r) {
                for (int c = 0; c < 3; ++c) Oi[r][c] = Od[c][r];
                Oi[r][3] = -(Od[0][r] * Od[0][3] + Od[1][r] * Od[1][3] + Od[2][r] * Od[2][3]);
            }
            Oi[3][3] = 1;
            for (size_t h = 0; h < handles.size(); ++h) {
                double p[4] = {xyz[3 * handles[h]], xyz[3 * handles[h] + 1], xyz[3 * handles[h] + 2], 1}, q[4], u[4], v[4];
                for (int r = 0; r < 4; ++r) q[r] = Oi[r][0] * p[0] + Oi[r][1] * p[1] + Oi[r][2] * p[2] + Oi[r][3] * p[3];
                for (int r = 0; r < 4; ++r) u[r] = Td[r][0] * q[0] + Td[r][1] * q[1] + Td[r][2] * q[2] + Td[r][3] * q[3];
                for (int r = 0; r < 4; ++r) v[r] = Od[r][0] * u[0] + Od[r][1] * u[1] + Od[r][2] * u[2] + Od[r][3] * u[3];
                for (int k = 0; k < 3; ++k) REQUIRE(std::fabs(arap.xyz[3 * h + k] - v[k]) < 1e-5);
            }
        }
        // at t = 0 the motion is the identity: the handles stay where they are
        Recorder arap0;
        dutil.updateConstraints(trajectory(0.f), arap0);
        for (size_t h = 0; h < handles.size(); ++h)
            for (int k = 0; k < 3; ++k) REQUIRE(std::fabs(arap0.xyz[3 * h + k] - xyz[3 * handles[h] + k]) < 1e-6);
    }
    std::printf(failures ? "transform types: FAILED\n" : "transform types: ok\n");
    return failures ? 1 : 0;
}
