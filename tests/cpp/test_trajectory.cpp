// tests/test_trajectory.cpp:17-39 with include/deform/trajectory.h (float, as the reference),
// then the same trajectory sampled in double for a comparison with the Python restatement.
#include <deform/trajectory.h>

#include <cmath>
#include <cstdio>

template <class S> deform::Transform<S> translation(S x, S y, S z) {
    auto t = deform::Transform<S>::Identity();
    t.m[0][3] = x; t.m[1][3] = y; t.m[2][3] = z;
    return t;
}

template <class S> deform::Transform<S> rotation_x(S a) {
    auto t = deform::Transform<S>::Identity();
    t.m[1][1] = std::cos(a); t.m[1][2] = -std::sin(a);
    t.m[2][1] = std::sin(a); t.m[2][2] = std::cos(a);
    return t;
}

template <class S> bool is_approx(const deform::Transform<S> &a, const deform::Transform<S> &b, S prec) {
    double d = 0, na = 0, nb = 0;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            d += (a.m[r][c] - b.m[r][c]) * (a.m[r][c] - b.m[r][c]);
            na += a.m[r][c] * a.m[r][c];
            nb += b.m[r][c] * b.m[r][c];
        }
    return std::sqrt(d) <= prec * std::sqrt(std::fmin(na, nb));
}

template <class S> deform::TrajectorySE3<S> reference_path(deform::Transform<S> poses[4]) {
    deform::TrajectorySE3<S> path;
    poses[0] = deform::Transform<S>::Identity();
    path.addKeyPose(poses[0]);
    poses[1] = translation<S>(0, 0, 1) * poses[0];
    path.addKeyPose(poses[1]);
    poses[2] = translation<S>(0, 0, 1) * rotation_x<S>((S)M_PI / 4) * poses[1];
    path.addKeyPose(poses[2]);
    poses[3] = translation<S>(0, 0, 1) * poses[2];
    path.addKeyPose(poses[3]);
    return path;
}

int main() {
    deform::Transform<float> pf[4];
    auto path = reference_path<float>(pf);
    bool ok0 = is_approx(path(0.f), pf[0], 1e-3f), ok1 = is_approx(path(1.f), pf[3], 1e-3f);
    printf("endpoints %d %d\n", ok0 ? 1 : 0, ok1 ? 1 : 0);
    deform::Transform<double> pd[4];
    auto pathd = reference_path<double>(pd);
    for (double t : {0.0, 0.2, 0.5, 0.8, 1.0}) {
        auto T = pathd(t);
        printf("t %.3f", t);
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) printf(" %.17g", T.m[r][c]);
        printf("\n");
    }
    return ok0 && ok1 ? 0 : 1;
}
