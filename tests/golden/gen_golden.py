"""Writes the golden fixtures under tests/golden/ (run in the build container).

Inputs come from the reference's data files, read as data:
  * etc/sphere.obj, etc/bar.obj, etc/plane.obj (meshes), etc/anchors.txt, etc/handles.txt
  * the anchor/handle index lists of examples/deform_bar.cpp:30-33 and the vertex indices
    37 (anchor) and 32 (handle) of examples/deform_sphere.cpp:60,65
  * the known answer of tests/test_cotan.cpp:20-54 (unit square).
Expected outputs come from two independent sources:
  * numpy.linalg.svd for the 3x3 rotations of arap.h:376-382 (not the oracle), and
  * the FP64 oracle (oracle/arap_oracle.c) for whole deformations, cross-checked here against
    an independent vectorised numpy/scipy ARAP (``numpy_arap`` below) before being written.
The fixtures are data only (arrays of inputs and expected outputs).

Usage: python tests/golden/gen_golden.py [/root/reference]
"""
import os
import re
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "mesh-deform_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import meshgen  # noqa: E402
import objio  # noqa: E402
import oracle  # noqa: E402


def numpy_rotation(cov):
    U, S, Vt = np.linalg.svd(cov)
    V = Vt.T
    D = np.eye(3)
    D[2, 2] = np.linalg.det(V @ U.T)
    return V @ D @ U.T


def numpy_arap(xyz, tri, con_idx, con_xyz, iters, weights=None):
    """Independent vectorised ARAP (numpy SVD + scipy sparse LU) used to cross-check the oracle."""
    n = xyz.shape[0]
    if weights is None:
        rp, ci, w = oracle.cotan_csr(xyz, tri)
    else:
        rp, ci, w = weights
    W = sp.csr_matrix((w, ci, rp), shape=(n, n))
    con = np.zeros(n, bool)
    con[con_idx] = True
    cpos = np.zeros((n, 3))
    cpos[con_idx] = con_xyz
    free = np.nonzero(~con)[0]
    deg = np.asarray(W.sum(axis=1)).ravel()
    Wff = W[free][:, free]
    L = (sp.diags(deg[free]) - Wff).tocsc()
    bfixed = W[free][:, np.nonzero(con)[0]] @ cpos[con]
    lu = spla.splu(L)
    rows = np.repeat(np.arange(n), np.diff(rp))
    e = xyz[rows] - xyz[ci]
    pp = xyz.copy()
    pp[con] = cpos[con]
    energies = []
    R = None
    for _ in range(iters):
        ep = pp[rows] - pp[ci]
        outer = w[:, None, None] * e[:, :, None] * ep[:, None, :]
        cov = np.zeros((n, 3, 3))
        np.add.at(cov, rows, outer)
        R = np.stack([numpy_rotation(c) for c in cov])
        rhs_e = 0.5 * w[:, None] * np.einsum("kab,kb->ka", R[rows] + R[ci], e)
        b = np.zeros((n, 3))
        np.add.at(b, rows, rhs_e)
        b = b[free] + bfixed
        x = lu.solve(b)
        pp[free] = x
        ep = pp[rows] - pp[ci]
        res = ep - np.einsum("kab,kb->ka", R[rows], e)
        energies.append(float(np.sum(w * np.sum(res * res, axis=1))))
    return pp, R, np.array(energies)


def oracle_case(xyz, tri, con_idx, con_xyz, iters):
    a = oracle.Arap(xyz, tri)
    a.set_constraints(con_idx, con_xyz)
    ok, en = a.deform(iters, energy=True)
    assert ok
    rp, ci, w = a.cotan()
    # rotations after the first iteration, from a fresh object
    a1 = oracle.Arap(xyz, tri)
    a1.set_constraints(con_idx, con_xyz)
    a1.deform(1)
    return dict(rp=rp, ci=ci, w=w, pos=a.mesh.copy(), R1=a1.rotations(), energy=en,
                factor_nnz=np.int64(a.factor_nnz()))


def check_against_numpy(name, xyz, tri, con_idx, con_xyz, iters, got):
    pp, _, en = numpy_arap(xyz, tri, con_idx, con_xyz, iters)
    err = meshgen.relative_error(got["pos"], pp)
    eerr = float(np.max(np.abs(got["energy"] - en) / np.maximum(np.abs(en), 1e-300)))
    print("%-14s n=%-6d iters=%-3d oracle vs numpy/scipy: pos rel %.2e  energy rel %.2e" %
          (name, xyz.shape[0], iters, err, eerr))
    assert err < 1e-10 and eerr < 1e-9, name


def int_list(src, name):
    m = re.search(r"%s\s*=\s*\{([^}]*)\}" % name, src)
    return np.array([int(t) for t in m.group(1).replace(" ", "").split(",") if t], np.int32)


def main(ref):
    out = {}
    # 1. known answer of tests/test_cotan.cpp:20-54
    sq_xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
    sq_tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    sq_expected = np.array([[0, .5, 0, .5], [.5, 0, .5, 0], [0, .5, 0, .5], [.5, 0, .5, 0]], np.float32)
    np.savez_compressed(os.path.join(HERE, "unit_square.npz"), xyz=sq_xyz, tri=sq_tri, expected=sq_expected)

    # 2. rotations from numpy.linalg.svd
    rng = np.random.default_rng(20161)
    covs = rng.standard_normal((256, 3, 3))
    covs[32:64] = covs[32:64] * np.array([1.0, 1.0, 1e-3])[None, None, :]              # near rank 2
    covs[64:96] = -covs[64:96]                                                            # det flips
    q = np.linalg.qr(rng.standard_normal((32, 3, 3)))[0]
    covs[96:128] = np.einsum("kij,kjl->kil", np.diag([3.0, 2.0, 1.0])[None], q)           # exact rotations
    Rs = np.stack([numpy_rotation(c) for c in covs])
    np.savez_compressed(os.path.join(HERE, "svd_rotations.npz"), cov=covs, R=Rs)

    # 3. reference example meshes
    sph_xyz, sph_tri = objio.read_obj(os.path.join(ref, "etc", "sphere.obj"))
    bar_xyz, bar_tri = objio.read_obj(os.path.join(ref, "etc", "bar.obj"))
    pl_xyz, pl_tri = objio.read_obj(os.path.join(ref, "etc", "plane.obj"))
    bar_src = open(os.path.join(ref, "examples", "deform_bar.cpp")).read()
    bar_handles = int_list(bar_src, "handles")
    bar_anchors = int_list(bar_src, "anchors")
    pl_anchors = np.array(open(os.path.join(ref, "etc", "anchors.txt")).read().split(), np.int32)
    pl_handles = np.array(open(os.path.join(ref, "etc", "handles.txt")).read().split(), np.int32)

    cases = []
    # sphere: anchor 37 at rest, handle 32 lifted by 0.5 (deform_sphere.cpp:60-91), 5 iterations
    cases.append(("sphere_obj", sph_xyz, sph_tri, np.array([37, 32], np.int32),
                  np.stack([sph_xyz[37], sph_xyz[32] + [0, 0, 0.5]]), 5))
    # bar: anchors at rest (deform_bar.cpp:70-73), handles rotated about x as applyRotationToHandles
    # does (deform_bar.cpp:47-53,77-78; here by 90 degrees in one step), 20 iterations (:94)
    rx = meshgen.rot_x(np.pi / 2)
    cases.append(("bar_obj", bar_xyz, bar_tri, np.concatenate([bar_anchors, bar_handles]),
                  np.concatenate([bar_xyz[bar_anchors], bar_xyz[bar_handles] @ rx.T]), 20))
    # plane (planar, rank-2 covariances): etc/anchors.txt at rest, etc/handles.txt lifted 0.3
    cases.append(("plane_obj", pl_xyz, pl_tri, np.concatenate([pl_anchors, pl_handles]),
                  np.concatenate([pl_xyz[pl_anchors], pl_xyz[pl_handles] + [0, 0, 0.3]]), 10))
    # the CPU-runnable UV sphere workload, 10 iterations
    pr = meshgen.sphere_problem()
    idx, tgt = pr.constraints()
    cases.append(("uv_sphere", pr.xyz, pr.tri, idx, tgt, 10))

    for name, xyz, tri, ci_, cx, iters in cases:
        got = oracle_case(xyz, tri, ci_, cx, iters)
        check_against_numpy(name, xyz, tri, ci_, cx, iters, got)
        np.savez_compressed(os.path.join(HERE, "%s.npz" % name), xyz=xyz, tri=tri, con_idx=ci_,
                            con_xyz=cx, iters=np.int32(iters), **got)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
