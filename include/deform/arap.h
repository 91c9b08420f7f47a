/*
 * deform/arap.h -- drop-in for inc/deform/arap.h of cheind/mesh-deform.
 *
 * Same class template, same member functions and the same mesh-adapter concept as the
 * reference (AsRigidAsPossibleDeformation<MeshType, PrecisionType>, inc/deform/arap.h:49-466);
 * the work runs on an MI355X through the C ABI of libmdarap.so (md_arap.h).
 *
 * The adapter concept is unchanged (inc/deform/openmesh_adapter.h:49-118):
 *   typedef ... Scalar;  typedef ... VertexType;   // VertexType(x, y, z) and v[i] or v(i)
 *   VertexType vertexLocation(int) const;  void vertexLocation(int, const VertexType&);
 *   FaceType face(int) const (3 vertex indices, f[i] or f(i));
 *   int numberOfFaces() const;  int numberOfVertices() const;
 *
 * PrecisionType = double runs every step in FP64; PrecisionType = float runs the local step
 * and the triangular solves in FP32 with an FP64 factor and FP64 residual refinement.
 *
 * Differences from the reference, all where it had undefined behaviour or no error path:
 *   - an out-of-range constraint index makes deform() return false (the reference writes
 *     out of bounds at arap.h:279);
 *   - a mesh component without constraints makes the system singular: deform() returns
 *     false (the reference's SimplicialLDLT may or may not notice, arap.h:337-339);
 *   - the constructor throws std::runtime_error when no MI355X is available.
 */
#ifndef DEFORM_ARAP_H
#define DEFORM_ARAP_H

#include <md_arap.h>

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

namespace deform {

template <class T> class PrivateAccessor;

/** Minimal 3-vector accepted by setConstraint when Eigen is not used (any type with
 *  operator[] or operator() works too, Eigen::Matrix<S,3,1> included). */
template <class S> struct Vector3 {
    S v[3];
    Vector3() : v{S(0), S(0), S(0)} {}
    Vector3(S x, S y, S z) : v{x, y, z} {}
    S &operator[](int i) { return v[i]; }
    const S &operator[](int i) const { return v[i]; }
    S &operator()(int i) { return v[i]; }
    const S &operator()(int i) const { return v[i]; }
};

namespace detail {

template <class V, class = void> struct has_subscript : std::false_type {};
template <class V>
struct has_subscript<V, decltype((void)std::declval<const V &>()[0], void())> : std::true_type {};

template <class V> double component(const V &v, int i, std::true_type) { return (double)v[i]; }
template <class V> double component(const V &v, int i, std::false_type) { return (double)v(i); }
template <class V> double component(const V &v, int i) { return component(v, i, has_subscript<V>()); }

/** One context (device 0, one HIP stream) shared by every ARAP object of the process. */
inline md_ctx *context() {
    struct Holder {
        md_ctx *ctx = nullptr;
        int rc = MD_OK;
        Holder() { rc = md_ctx_create(0, &ctx); }
        ~Holder() { if (ctx) md_ctx_destroy(ctx); }
    };
    static Holder h;
    if (h.rc != MD_OK) throw std::runtime_error(std::string("mesh-deform: ") + md_last_error());
    return h.ctx;
}

}  // namespace detail

/**
 * As-rigid-as-possible surface deformation (Sorkine and Alexa, SGP 2007): the interface of
 * inc/deform/arap.h:49-138.
 */
template <class MeshType, class PrecisionType = typename MeshType::Scalar>
class AsRigidAsPossibleDeformation {
public:
    /** Scalar type of the positions and targets the class takes and returns. */
    typedef PrecisionType Scalar;
    /** Adapter the class reads vertices and faces through. */
    typedef MeshType Mesh;
    /** Vertex index type. */
    typedef int Index;

    /** Construct from mesh (arap.h:66-70). The mesh must outlive this object. */
    explicit AsRigidAsPossibleDeformation(Mesh &mesh) : _mesh(mesh), _handle(nullptr), _dirty(true) {
        const int n = _mesh.numberOfVertices(), F = _mesh.numberOfFaces();
        std::vector<int> tri(3 * (size_t)F);
        for (int f = 0; f < F; ++f) {  // initializeMeshTopology, arap.h:149-155
            auto fc = _mesh.face(f);
            for (int k = 0; k < 3; ++k) tri[3 * (size_t)f + k] = (int)detail::component(fc, k);
        }
        md_config cfg;
        md_config_default(&cfg);
        cfg.precision = std::is_same<PrecisionType, float>::value ? MD_PREC_FP32_REFINE : MD_PREC_FP64;
        cfg.rest_mode = MD_REST_REFERENCE;
        if (md_mesh_create(detail::context(), n, F, tri.data(), &cfg, &_handle) != MD_OK)
            throw std::runtime_error(std::string("mesh-deform: ") + md_last_error());
        _xyz.resize(3 * (size_t)n);
    }

    ~AsRigidAsPossibleDeformation() {
        if (_handle) md_mesh_destroy(_handle);
    }

    AsRigidAsPossibleDeformation(const AsRigidAsPossibleDeformation &) = delete;
    AsRigidAsPossibleDeformation &operator=(const AsRigidAsPossibleDeformation &) = delete;

    /** Pin a vertex to a location (arap.h:81-85); marks the object dirty. */
    template <class V> void setConstraint(Index vidx, const V &loc) {
        std::array<Scalar, 3> p = {(Scalar)detail::component(loc, 0), (Scalar)detail::component(loc, 1),
                                   (Scalar)detail::component(loc, 2)};
        _constrainedLocations[vidx] = p;
        _dirty = true;
    }

    /**
     * Run numberOfIterations local/global iterations and write the result back to the mesh
     * (arap.h:101-138). On a dirty call the current mesh becomes the rest pose and the initial
     * guess. Returns false when the system could not be factored.
     */
    bool deform(Index numberOfIterations) {
        const int n = _mesh.numberOfVertices();
        if (_dirty) {
            for (int i = 0; i < n; ++i) {  // initializeMeshGeometry, arap.h:162-168
                auto v = _mesh.vertexLocation(i);
                for (int k = 0; k < 3; ++k) _xyz[3 * (size_t)i + k] = (double)(Scalar)detail::component(v, k);
            }
            std::vector<int> idx;
            std::vector<double> xyz;
            idx.reserve(_constrainedLocations.size());
            xyz.reserve(3 * _constrainedLocations.size());
            for (const auto &c : _constrainedLocations) {
                idx.push_back(c.first);
                for (int k = 0; k < 3; ++k) xyz.push_back((double)c.second[k]);
            }
            if (md_mesh_set_constraints(_handle, (int)idx.size(), idx.data(), xyz.data()) != MD_OK) return false;
            if (_constrainedLocations.empty()) {
                int ok = 0;  // computes the weights, returns true without write-back (arap.h:113-114)
                return md_mesh_deform(_handle, 0, _xyz.data(), &ok) == MD_OK && ok;
            }
        }
        int ok = 0;
        int rc = md_mesh_deform(_handle, numberOfIterations, _xyz.data(), &ok);
        if (rc != MD_OK || !ok) return false;
        _dirty = false;
        typedef typename Mesh::Scalar MS;
        for (int i = 0; i < n; ++i)  // write-back, arap.h:133-135
            _mesh.vertexLocation(i, typename Mesh::VertexType((MS)_xyz[3 * (size_t)i], (MS)_xyz[3 * (size_t)i + 1],
                                                              (MS)_xyz[3 * (size_t)i + 2]));
        return true;
    }

    /** The C ABI handle, for callers that batch this mesh (md_batch_*). A server (md_server_*)
     *  takes meshes in fixed rest mode only; this class creates them in reference rest mode. */
    md_mesh *handle() const { return _handle; }

private:
    template <class> friend class PrivateAccessor;

    /** Cotan weight matrix of the current rest pose as CSR (the reference's _edgeWeights). */
    bool cotanWeights(std::vector<int> &rowptr, std::vector<int> &col, std::vector<double> &val) const {
        int64_t nnz = 0;
        if (md_mesh_cotan_nnz(_handle, &nnz) != MD_OK) return false;
        rowptr.assign((size_t)_mesh.numberOfVertices() + 1, 0);
        col.assign((size_t)nnz, 0);
        val.assign((size_t)nnz, 0.0);
        return md_mesh_cotan_csr(_handle, rowptr.data(), col.data(), val.data()) == MD_OK;
    }

    Mesh &_mesh;
    md_mesh *_handle;
    std::unordered_map<Index, std::array<Scalar, 3>> _constrainedLocations;
    std::vector<double> _xyz;
    bool _dirty;
};

}  // namespace deform

#endif
