/*
 * deform/deformation_util.h -- counterpart of inc/deform/deformation_util.h of
 * cheind/mesh-deform (DeformationUtil, deformation_util.h:19-64): captures a handle set and
 * moves it rigidly, conjugated by an origin frame, through setConstraint.
 */
#ifndef DEFORM_DEFORMATION_UTIL_H
#define DEFORM_DEFORMATION_UTIL_H

#include <deform/arap.h>
#include <deform/trajectory.h>

#include <array>
#include <vector>

namespace deform {

template <class MeshType> class DeformationUtil {
public:
    typedef MeshType Mesh;
    /** Floating point precision used in calculations. */
    typedef typename MeshType::Scalar Scalar;
    /** Transformation matrix type. */
    typedef deform::Transform<Scalar> Transform;

    /** Captures the initial handle positions (deformation_util.h:34-46). */
    template <class HandleIterator>
    DeformationUtil(const Mesh &mesh, HandleIterator handlesBegin, HandleIterator handlesEnd,
                    const Transform &origin = Transform::Identity())
        : _origin(origin), _originInv(origin.isometryInverse()) {
        _handles.insert(_handles.begin(), handlesBegin, handlesEnd);
        _points.resize(_handles.size());
        for (size_t i = 0; i < _handles.size(); ++i) {
            auto v = mesh.vertexLocation(_handles[i]);
            for (int k = 0; k < 3; ++k) _points[i][k] = (Scalar)detail::component(v, k);
        }
    }

    /** setConstraint(h, origin * t * origin^-1 * p0_h) for every handle (deformation_util.h:48-57). */
    template <class ARAP> void updateConstraints(const Transform &t, ARAP &arap) {
        Transform tabs = _origin * t * _originInv;
        for (size_t i = 0; i < _handles.size(); ++i) {
            std::array<Scalar, 3> p = tabs.apply(_points[i]);
            arap.setConstraint(_handles[i], Vector3<Scalar>(p[0], p[1], p[2]));
        }
    }

private:
    std::vector<int> _handles;
    std::vector<std::array<Scalar, 3>> _points;
    Transform _origin;
    Transform _originInv;
};

}  // namespace deform

#endif
