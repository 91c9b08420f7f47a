"""Minimal Wavefront OBJ reader for triangle meshes.

Reads ``v x y z`` records and ``f`` records in the ``a``, ``a/t``, ``a//n`` and ``a/t/n``
forms (1-based, negative indices relative to the end), which covers the reference's example
meshes etc/sphere.obj (``f a b c``), etc/bar.obj and etc/plane.obj (``f a//n ...``). Polygons
with more than three corners are fan-triangulated. Vertex order is file order, as the
reference's OpenMesh loader keeps it (examples/deform_sphere.cpp:38-44 then indexes
vertices 32 and 37 by that order).
"""
import numpy as np


def parse_obj(text):
    verts = []
    faces = []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
        elif parts[0] == "f":
            idx = []
            for tok in parts[1:]:
                k = int(tok.split("/")[0])
                idx.append(k - 1 if k > 0 else len(verts) + k)
            for j in range(1, len(idx) - 1):
                faces.append([idx[0], idx[j], idx[j + 1]])
    xyz = np.asarray(verts, np.float64).reshape(-1, 3)
    tri = np.asarray(faces, np.int32).reshape(-1, 3)
    if tri.size and (tri.min() < 0 or tri.max() >= xyz.shape[0]):
        raise ValueError("face index out of range")
    return xyz, tri


def read_obj(path):
    with open(path, "r") as f:
        return parse_obj(f.read())


def write_obj(path, xyz, tri):
    with open(path, "w") as f:
        for v in np.asarray(xyz, np.float64):
            f.write("v %.17g %.17g %.17g\n" % (v[0], v[1], v[2]))
        for t in np.asarray(tri, np.int64):
            f.write("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))
