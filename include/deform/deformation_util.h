/*
 * deform/deformation_util.h -- counterpart of inc/deform/deformation_util.h of
 * cheind/mesh-deform (DeformationUtil, deformation_util.h:19-64): captures a handle set and
 * moves it rigidly, conjugated by an origin frame, through setConstraint.
 *
 * Transforms are of the caller's type, as in trajectory.h: the class's Transform typedef is
 * Eigen::Transform<Scalar, 3, Affine> where Eigen is available (the reference's type) and
 * deform::Transform<Scalar> otherwise, and the constructor and updateConstraints accept any
 * type that reads as a 4x4 affine matrix (a TrajectorySE3<float>::Transform passed to a
 * DeformationUtil of a double mesh included, as examples/deform_trajectory.cpp:53-56,84 do).
 */
#ifndef DEFORM_DEFORMATION_UTIL_H
#define DEFORM_DEFORMATION_UTIL_H

#include <deform/arap.h>
#include <deform/trajectory.h>

#include <array>
#include <vector>

namespace deform {

template <class MeshType, class TransformType = DefaultTransform<typename MeshType::Scalar>>
class DeformationUtil {
public:
    typedef MeshType Mesh;
    /** Scalar type of the mesh; the handle motion is computed in it. */
    typedef typename MeshType::Scalar Scalar;
    /** Rigid motion type of the interface (any 4x4-affine-readable type is accepted). */
    typedef TransformType Transform;

    /** Records the handles and their current positions p0; origin is the frame the motions
     *  are expressed in (deformation_util.h:34-46). */
    template <class HandleIterator, class T = Transform>
    DeformationUtil(const Mesh &mesh, HandleIterator handlesBegin, HandleIterator handlesEnd,
                    const T &origin = T::Identity())
        : _frame(detail::to_rigid<Scalar>(origin)), _frameInv(_frame.isometryInverse()) {
        _ids.insert(_ids.begin(), handlesBegin, handlesEnd);
        _rest.resize(_ids.size());
        for (size_t i = 0; i < _ids.size(); ++i) {
            auto v = mesh.vertexLocation(_ids[i]);
            for (int k = 0; k < 3; ++k) _rest[i][k] = (Scalar)detail::component(v, k);
        }
    }

    /** Moves every handle to origin * t * origin^-1 * p0 through arap.setConstraint
     *  (deformation_util.h:48-57). */
    template <class T, class ARAP> void updateConstraints(const T &t, ARAP &arap) {
        const deform::Transform<Scalar> tabs = _frame * detail::to_rigid<Scalar>(t) * _frameInv;
        for (size_t i = 0; i < _ids.size(); ++i) {
            std::array<Scalar, 3> p = tabs.apply(_rest[i]);
            arap.setConstraint(_ids[i], Vector3<Scalar>(p[0], p[1], p[2]));
        }
    }

private:
    std::vector<int> _ids;
    std::vector<std::array<Scalar, 3>> _rest;
    deform::Transform<Scalar> _frame, _frameInv;
};

}  // namespace deform

#endif
