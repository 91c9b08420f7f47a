// Reads an OBJ file with include/deform/obj_io.h and prints "V F", then every vertex
// (17 significant digits) and every triangle (0-based), one per line; exit status 1 with the
// reader's message on a malformed file.
#include <deform/obj_io.h>

#include <cstdio>

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: test_objio FILE.obj\n");
        return 2;
    }
    std::vector<double> xyz;
    std::vector<int> tri;
    std::string err;
    if (!deform::readObj(argv[1], xyz, tri, &err)) {
        std::printf("error: %s\n", err.c_str());
        return 1;
    }
    std::printf("%zu %zu\n", xyz.size() / 3, tri.size() / 3);
    for (size_t i = 0; i < xyz.size(); i += 3) std::printf("%.17g %.17g %.17g\n", xyz[i], xyz[i + 1], xyz[i + 2]);
    for (size_t i = 0; i < tri.size(); i += 3) std::printf("%d %d %d\n", tri[i], tri[i + 1], tri[i + 2]);
    return 0;
}
