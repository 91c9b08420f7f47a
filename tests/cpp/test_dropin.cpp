// Drives include/deform/arap.h exactly as the reference's callers do:
//   cotan   -- tests/test_cotan.cpp:20-54 (unit square, deform(0), weights via PrivateAccessor)
//   sphere  -- one deform(iters) on a UV sphere (anchors z < -0.8, handles z > 0.8 moved)
//   frames  -- examples/deform_sphere.cpp:71-91 (setConstraint + deform(5) every frame)
// Results are printed as text; tests/test_dropin_cpp.py compares them with the oracle.
#include <deform/arap.h>
#include <deform/array_adapter.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace deform {
// White-box hook, the counterpart of tests/accessor.h:16-22.
template <class ARAP> class PrivateAccessor {
public:
    static bool cotanWeights(const ARAP &a, std::vector<int> &rp, std::vector<int> &ci, std::vector<double> &v) {
        return a.cotanWeights(rp, ci, v);
    }
};
}  // namespace deform

static void uv_sphere(int rings, int segs, std::vector<double> &xyz, std::vector<int> &tri) {
    xyz.clear();
    tri.clear();
    xyz.insert(xyz.end(), {0.0, 0.0, 1.0});
    for (int k = 1; k <= rings; ++k) {
        double th = k * M_PI / (rings + 1);
        for (int j = 0; j < segs; ++j) {
            double ph = j * 2.0 * M_PI / segs;
            xyz.insert(xyz.end(), {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)});
        }
    }
    xyz.insert(xyz.end(), {0.0, 0.0, -1.0});
    int south = (int)xyz.size() / 3 - 1;
    auto vid = [&](int k, int j) { return 1 + k * segs + (j % segs); };
    for (int j = 0; j < segs; ++j) tri.insert(tri.end(), {0, vid(0, j), vid(0, j + 1)});
    for (int k = 0; k + 1 < rings; ++k)
        for (int j = 0; j < segs; ++j) {
            int a = vid(k, j), b = vid(k, j + 1), c = vid(k + 1, j), d = vid(k + 1, j + 1);
            tri.insert(tri.end(), {a, c, d});
            tri.insert(tri.end(), {a, d, b});
        }
    for (int j = 0; j < segs; ++j) tri.insert(tri.end(), {south, vid(rings - 1, j + 1), vid(rings - 1, j)});
}

static int run_cotan() {
    // tests/test_cotan.cpp:25-41 with a float mesh (OpenMesh's default scalar)
    std::vector<float> xyz = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0};
    std::vector<int> tri = {0, 1, 2, 0, 2, 3};
    typedef deform::ArrayAdapter<float> Adapter;
    typedef deform::AsRigidAsPossibleDeformation<Adapter> ARAP;
    Adapter adapter(xyz, tri);
    ARAP arap(adapter);
    if (!arap.deform(0)) return 2;
    std::vector<int> rp, ci;
    std::vector<double> v;
    if (!deform::PrivateAccessor<ARAP>::cotanWeights(arap, rp, ci, v)) return 3;
    double W[4][4] = {};
    for (int i = 0; i < 4; ++i)
        for (int e = rp[i]; e < rp[i + 1]; ++e) W[i][ci[e]] = v[e];
    const float expected[4][4] = {{0, .5f, 0, .5f}, {.5f, 0, .5f, 0}, {0, .5f, 0, .5f}, {.5f, 0, .5f, 0}};
    double d2 = 0, n1 = 0, n2 = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float w = (float)W[i][j];
            d2 += (w - expected[i][j]) * (w - expected[i][j]);
            n1 += w * w;
            n2 += expected[i][j] * expected[i][j];
        }
    bool ok = std::sqrt(d2) <= 1e-4 * std::sqrt(std::fmin(n1, n2));  // Eigen isApprox(.., 1e-4f)
    for (int i = 0; i < 4; ++i) printf("%.17g %.17g %.17g %.17g\n", W[i][0], W[i][1], W[i][2], W[i][3]);
    printf("isApprox %d\n", ok ? 1 : 0);
    return ok ? 0 : 1;
}

template <class P> static int run_sphere(int rings, int segs, int iters) {
    std::vector<double> xyz;
    std::vector<int> tri;
    uv_sphere(rings, segs, xyz, tri);
    typedef deform::ArrayAdapter<double> Adapter;
    Adapter adapter(xyz, tri);
    deform::AsRigidAsPossibleDeformation<Adapter, P> arap(adapter);
    int n = (int)xyz.size() / 3;
    for (int i = 0; i < n; ++i) {
        double z = xyz[3 * i + 2];
        if (z < -0.8) arap.setConstraint(i, deform::Vector3<double>(xyz[3 * i], xyz[3 * i + 1], z));
        else if (z > 0.8) arap.setConstraint(i, deform::Vector3<double>(xyz[3 * i] + 0.3, xyz[3 * i + 1], z + 0.5));
    }
    if (!arap.deform(iters)) return 2;
    for (int i = 0; i < n; ++i) printf("%.17g %.17g %.17g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return 0;
}

static int run_frames(int rings, int segs, int frames) {
    std::vector<double> xyz;
    std::vector<int> tri;
    uv_sphere(rings, segs, xyz, tri);
    typedef deform::ArrayAdapter<double> Adapter;
    Adapter adapter(xyz, tri);
    deform::AsRigidAsPossibleDeformation<Adapter, double> arap(adapter);
    int n = (int)xyz.size() / 3;
    int anchor = n - 1, handle = 0;  // south pole pinned, north pole dragged
    arap.setConstraint(anchor, adapter.vertexLocation(anchor));
    double p[3] = {xyz[0], xyz[1], xyz[2]};
    double ph = -M_PI / 2;
    for (int f = 0; f < frames; ++f) {
        ph += 0.3;
        arap.setConstraint(handle, deform::Vector3<double>(p[0], p[1], p[2] + std::sin(ph)));
        if (!arap.deform(5)) return 2;
    }
    for (int i = 0; i < n; ++i) printf("%.17g %.17g %.17g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) return 64;
    try {
        if (!std::strcmp(argv[1], "cotan")) return run_cotan();
        if (!std::strcmp(argv[1], "sphere") && argc == 5)
            return run_sphere<double>(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
        if (!std::strcmp(argv[1], "sphere32") && argc == 5)
            return run_sphere<float>(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
        if (!std::strcmp(argv[1], "frames") && argc == 5) return run_frames(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]));
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 70;
    }
    return 64;
}
