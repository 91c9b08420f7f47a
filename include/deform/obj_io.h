/*
 * deform/obj_io.h -- Wavefront OBJ reading and writing for triangle meshes, for callers that
 * load the reference's example meshes (etc/sphere.obj, etc/bar.obj, etc/plane.obj) without
 * OpenMesh (the reference reads them with OpenMesh::IO::read_mesh, examples/deform_bar.cpp:61,
 * examples/deform_sphere.cpp:38-44).
 *
 * Reads `v x y z` records and `f` records whose corners are `a`, `a/t`, `a//n` or `a/t/n`
 * (1-based; negative indices count back from the last vertex read). Polygons with more than
 * three corners are split into a fan. Every other record (vn, vt, o, g, s, mtllib, usemtl,
 * comments) is skipped. Vertices keep file order, as OpenMesh's loader keeps it, so the
 * examples' vertex indices (handles and anchors) address the same vertices.
 */
#ifndef DEFORM_OBJ_IO_H
#define DEFORM_OBJ_IO_H

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace deform {

/** Parses OBJ text: 3 scalars per vertex into xyz, 3 vertex indices (0-based) per triangle
 *  into tri. Returns false with a message in err (when given) on a malformed record or a face
 *  index outside the vertices read so far. */
template <class S>
bool parseObj(std::istream &in, std::vector<S> &xyz, std::vector<int> &tri, std::string *err = nullptr) {
    xyz.clear();
    tri.clear();
    std::string line;
    long lineno = 0;
    auto fail = [&](const std::string &what) {
        if (err) *err = "line " + std::to_string(lineno) + ": " + what;
        return false;
    };
    std::vector<int> corners;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            double x, y, z;
            if (!(ls >> x >> y >> z)) return fail("a vertex record needs three coordinates");
            xyz.push_back((S)x);
            xyz.push_back((S)y);
            xyz.push_back((S)z);
        } else if (tag == "f") {
            corners.clear();
            const long nv = (long)(xyz.size() / 3);
            std::string tok;
            while (ls >> tok) {
                char *end = nullptr;
                const long k = std::strtol(tok.c_str(), &end, 10);
                if (end == tok.c_str() || (*end != '\0' && *end != '/')) return fail("bad face corner '" + tok + "'");
                const long v = k > 0 ? k - 1 : nv + k;
                if (k == 0 || v < 0 || v >= nv) return fail("face corner '" + tok + "' outside the vertices read");
                corners.push_back((int)v);
            }
            if (corners.size() < 3) return fail("a face needs three corners");
            for (size_t j = 1; j + 1 < corners.size(); ++j) {
                tri.push_back(corners[0]);
                tri.push_back(corners[j]);
                tri.push_back(corners[j + 1]);
            }
        }
    }
    return true;
}

/** Reads an OBJ file (see parseObj). */
template <class S>
bool readObj(const std::string &path, std::vector<S> &xyz, std::vector<int> &tri, std::string *err = nullptr) {
    std::ifstream f(path);
    if (!f) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    return parseObj(f, xyz, tri, err);
}

/** Writes `v` and `f a b c` records (positions with 17 significant digits: they read back
 *  bit for bit in double precision). */
template <class S> bool writeObj(const std::string &path, const std::vector<S> &xyz, const std::vector<int> &tri) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    for (size_t i = 0; i + 2 < xyz.size(); i += 3)
        std::fprintf(f, "v %.17g %.17g %.17g\n", (double)xyz[i], (double)xyz[i + 1], (double)xyz[i + 2]);
    for (size_t i = 0; i + 2 < tri.size(); i += 3) std::fprintf(f, "f %d %d %d\n", tri[i] + 1, tri[i + 1] + 1, tri[i + 2] + 1);
    return std::fclose(f) == 0;
}

}  // namespace deform

#endif
