/*
 * deform/array_adapter.h -- mesh adapter over plain arrays.
 *
 * Provides the five members the ARAP class needs (the adapter concept of
 * inc/deform/openmesh_adapter.h:49-118 of cheind/mesh-deform) for a mesh held as a vertex
 * array (3 scalars per vertex) and a triangle array (3 vertex indices per face, counter-
 * clockwise). The arrays are referenced, not copied, and must outlive the adapter.
 */
#ifndef DEFORM_ARRAY_ADAPTER_H
#define DEFORM_ARRAY_ADAPTER_H

#include <deform/arap.h>

#include <vector>

namespace deform {

template <class S = double> class ArrayAdapter {
public:
    /** Scalar type of the stored positions (concept member). */
    typedef S Scalar;
    typedef Vector3<S> VertexType;
    typedef Vector3<int> FaceType;

    ArrayAdapter(std::vector<S> &xyz, const std::vector<int> &tri) : _xyz(xyz), _tri(tri) {}

    /** Position of vertex idx (concept member). */
    VertexType vertexLocation(int idx) const {
        return VertexType(_xyz[3 * (size_t)idx], _xyz[3 * (size_t)idx + 1], _xyz[3 * (size_t)idx + 2]);
    }
    /** Overwrites the position of vertex idx (concept member). */
    void vertexLocation(int idx, const VertexType &v) {
        _xyz[3 * (size_t)idx] = v[0];
        _xyz[3 * (size_t)idx + 1] = v[1];
        _xyz[3 * (size_t)idx + 2] = v[2];
    }
    /** The three vertex indices of face idx, counter-clockwise (concept member). */
    FaceType face(int idx) const {
        return FaceType(_tri[3 * (size_t)idx], _tri[3 * (size_t)idx + 1], _tri[3 * (size_t)idx + 2]);
    }
    /** Face count (concept member). */
    int numberOfFaces() const { return (int)(_tri.size() / 3); }
    /** Vertex count (concept member). */
    int numberOfVertices() const { return (int)(_xyz.size() / 3); }

private:
    std::vector<S> &_xyz;
    const std::vector<int> &_tri;
};

}  // namespace deform

#endif
