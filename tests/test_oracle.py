"""Pins the FP64 oracle (oracle/arap_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement of inc/deform/arap.h
fails it: the reference's own known answer (tests/test_cotan.cpp:20-54), closed forms of the
cotangent clamps (arap.h:199-218), numpy.linalg.svd rotations (arap.h:376-382), scipy sparse
solves (arap.h:337,420), rigid-motion exactness, optimality of each half-step and the
semantics of deform() (arap.h:101-138).
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import meshgen
import oracle


def dense(rp, ci, v, n):
    return sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()


# --------------------------------------------------------------------------------------
# cotan weights (arap.h:182-239)
# --------------------------------------------------------------------------------------

def test_cotan_reference_known_answer(golden_dir):
    """tests/test_cotan.cpp:20-54: unit square split on 0-2 -> 0.5 on the boundary edges."""
    g = np.load(os.path.join(golden_dir, "unit_square.npz"))
    rp, ci, v = oracle.cotan_csr(g["xyz"], g["tri"])
    W = dense(rp, ci, v, 4).astype(np.float32)
    exp = g["expected"]
    # Eigen isApprox(expected, 1e-4): ||W - E|| <= 1e-4 * min(||W||, ||E||)
    assert np.linalg.norm(W - exp) <= 1e-4 * min(np.linalg.norm(W), np.linalg.norm(exp))
    # the diagonal 0-2 carries the two clamped right-angle cotangents: 2 * 0.5 * 1e-10
    assert dense(rp, ci, v, 4)[0, 2] == pytest.approx(1e-10, rel=1e-12)


def test_cotan_equilateral_closed_form():
    s3 = math.sqrt(3.0)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0.5, s3 / 2, 0]], np.float64)
    rp, ci, v = oracle.cotan_csr(xyz, np.array([[0, 1, 2]], np.int32))
    assert len(v) == 6
    np.testing.assert_allclose(v, 0.5 / math.tan(math.pi / 3), rtol=1e-14)


def test_cotan_obtuse_and_degenerate_clamps():
    # obtuse angle at vertex 2 (opposite edge 0-1): cot < 0 -> clamped to 1e-10
    xyz = np.array([[0, 0, 0], [2, 0, 0], [1, 0.2, 0]], np.float64)
    rp, ci, v = oracle.cotan_csr(xyz, np.array([[0, 1, 2]], np.int32))
    W = dense(rp, ci, v, 3)
    assert W[0, 1] == pytest.approx(0.5e-10, rel=1e-12)
    # acute corners keep cot = adjacent/opposite
    assert W[1, 2] == pytest.approx(0.5 * (1.0 / 0.2), rel=1e-12)
    # coincident vertices: squared lengths clamp to 1e-8, area clamps to 1e-8 -> cot = 0.25
    xyz = np.zeros((3, 3))
    rp, ci, v = oracle.cotan_csr(xyz, np.array([[0, 1, 2]], np.int32))
    np.testing.assert_allclose(v, 0.125, rtol=1e-6)
    # NaN coordinates: std::max(1e-8, NaN) returns the clamp (arap.h:199-208 argument order)
    xyz = np.full((3, 3), np.nan)
    rp, ci, v = oracle.cotan_csr(xyz, np.array([[0, 1, 2]], np.int32))
    assert np.all(np.isfinite(v))
    np.testing.assert_allclose(v, 0.125, rtol=1e-6)


def test_cotan_structure_on_workload():
    pr = meshgen.sphere_problem()
    rp, ci, v = oracle.cotan_csr(pr.xyz, pr.tri)
    n = pr.n
    E = pr.F * 3 // 2  # closed manifold
    assert len(v) == 2 * E
    W = sp.csr_matrix((v, ci, rp), shape=(n, n))
    assert abs(W - W.T).max() == 0.0
    assert np.all(v >= 1e-10 * 0.5)
    for i in range(0, n, 97):  # ascending columns, no diagonal
        cols = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0) and i not in cols
    # duplicate triplets are summed: interior edge weight = 0.5 (cot a + cot b)
    # (checked against an explicit per-face sum)
    acc = {}
    x = pr.xyz
    for f in pr.tri[:200]:
        for k in range(3):
            a, b, c = f[k], f[(k + 1) % 3], f[(k + 2) % 3]
            u, w_ = x[a] - x[c], x[b] - x[c]
            cot = np.dot(u, w_) / np.linalg.norm(np.cross(u, w_))
            acc.setdefault((min(a, b), max(a, b)), []).append(max(1e-10, cot))
    Wd = W.tocsr()
    for (a, b), cots in acc.items():
        if len(cots) == 2:
            assert Wd[a, b] == pytest.approx(0.5 * sum(cots), rel=1e-9)


# --------------------------------------------------------------------------------------
# rotations (arap.h:354-384)
# --------------------------------------------------------------------------------------

def test_rotations_match_numpy_svd(golden_dir):
    g = np.load(os.path.join(golden_dir, "svd_rotations.npz"))
    for cov, R in zip(g["cov"], g["R"]):
        Ro = oracle.svd_rotation(cov)
        np.testing.assert_allclose(Ro, R, atol=1e-10)
        np.testing.assert_allclose(Ro @ Ro.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(Ro) == pytest.approx(1.0, abs=1e-12)


def test_svd_descending_and_reconstructs():
    rng = np.random.default_rng(7)
    for _ in range(100):
        A = rng.standard_normal((3, 3))
        U, S, V = oracle.svd3(A)
        assert S[0] >= S[1] >= S[2] >= 0
        np.testing.assert_allclose(U @ np.diag(S) @ V.T, A, atol=1e-12)
        np.testing.assert_allclose(S, np.linalg.svd(A, compute_uv=False), rtol=1e-12, atol=1e-14)


def test_rotation_maximises_trace():
    """R = argmax_{SO(3)} tr(R cov): no random rotation does better."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        cov = rng.standard_normal((3, 3))
        R = oracle.svd_rotation(cov)
        best = np.trace(R @ cov)
        q = np.linalg.qr(rng.standard_normal((500, 3, 3)))[0]
        q[np.linalg.det(q) < 0, :, 0] *= -1
        assert np.all(np.einsum("kij,ji->k", q, cov) <= best + 1e-12)


def test_rotations_rigid_and_identity():
    pr = meshgen.sphere_problem()
    rp, ci, w = oracle.cotan_csr(pr.xyz, pr.tri)
    R = oracle.rotations(rp, ci, w, pr.xyz, pr.xyz)
    np.testing.assert_allclose(R, np.broadcast_to(np.eye(3), R.shape), atol=1e-12)
    Q = np.linalg.qr(np.random.default_rng(3).standard_normal((3, 3)))[0]
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    R = oracle.rotations(rp, ci, w, pr.xyz, pr.xyz @ Q.T + [1.0, -2.0, 0.5])
    np.testing.assert_allclose(R, np.broadcast_to(Q, R.shape), atol=1e-11)


# --------------------------------------------------------------------------------------
# sparse LDL^T (arap.h:337,420)
# --------------------------------------------------------------------------------------

def _laplacian(pr):
    n = pr.n
    rp, ci, w = oracle.cotan_csr(pr.xyz, pr.tri)
    W = sp.csr_matrix((w, ci, rp), shape=(n, n))
    con = np.zeros(n, bool)
    con[pr.anchors] = True
    con[pr.handles] = True
    free = np.nonzero(~con)[0]
    deg = np.asarray(W.sum(axis=1)).ravel()
    L = (sp.diags(deg[free]) - W[free][:, free]).tocsr()
    L.sort_indices()
    return L, pr.xyz[free]


def test_ldl_solve_matches_scipy():
    pr = meshgen.sphere_problem()
    L, fx = _laplacian(pr)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((3, L.shape[0]))
    x, nnzl = oracle.spd_solve(L.indptr, L.indices, L.data, b, xyz=fx)
    ref = spla.spsolve(L.tocsc(), b.T).T
    np.testing.assert_allclose(x, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert np.max(np.abs(L @ x.T - b.T)) < 1e-10 * np.abs(b).max()
    # natural order gives the same solution, with more fill
    x2, nnzl2 = oracle.spd_solve(L.indptr, L.indices, L.data, b)
    np.testing.assert_allclose(x2, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    assert nnzl < nnzl2


def test_nd_order_is_permutation():
    pr = meshgen.bar_problem(40, 4, 4)
    L, fx = _laplacian(pr)
    perm = np.zeros(L.shape[0], np.int32)
    assert oracle.lib().or_nd_order(L.shape[0], L.indptr.astype(np.int32), L.indices.astype(np.int32),
                                    np.ascontiguousarray(fx).ravel(), perm) == 0
    assert np.array_equal(np.sort(perm), np.arange(L.shape[0]))


# --------------------------------------------------------------------------------------
# whole deform() (arap.h:101-138)
# --------------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "plane_obj", "uv_sphere"])
def test_golden_deformations(golden_dir, name):
    g = np.load(os.path.join(golden_dir, name + ".npz"))
    a = oracle.Arap(g["xyz"], g["tri"])
    a.set_constraints(g["con_idx"], g["con_xyz"])
    ok, en = a.deform(int(g["iters"]), energy=True)
    assert ok
    np.testing.assert_allclose(a.mesh, g["pos"], rtol=0, atol=1e-12 * np.abs(g["pos"]).max())
    np.testing.assert_allclose(en, g["energy"], rtol=1e-12)
    rp, ci, w = a.cotan()
    assert np.array_equal(rp, g["rp"]) and np.array_equal(ci, g["ci"])
    np.testing.assert_allclose(w, g["w"], rtol=1e-15)
    # energy is non-increasing: both half-steps are exact minimisers
    assert np.all(np.diff(en) <= 1e-12 * en[0])


def test_global_step_optimality():
    """After deform, L p'_free = b (the rhs of arap.h:395-414 with the last rotations)."""
    pr = meshgen.sphere_problem()
    a = oracle.Arap(pr.xyz, pr.tri)
    idx, tgt = pr.constraints()
    a.set_constraints(idx, tgt)
    a.deform(3)
    L, _ = _laplacian(pr)
    free = np.setdiff1d(np.arange(pr.n), idx)
    b = a.rhs()
    r = L @ a.pprime()[free] - b
    assert np.abs(r).max() < 1e-11 * np.abs(b).max()
    # rhs closed form with the last rotations: bFixed + sum_j w/2 (R_i + R_j)(p_i - p_j)
    rp, ci, w = a.cotan()
    R = a.rotations()
    rows = np.repeat(np.arange(pr.n), np.diff(rp))
    e = pr.xyz[rows] - pr.xyz[ci]
    contrib = 0.5 * w[:, None] * np.einsum("kab,kb->ka", R[rows] + R[ci], e)
    full = np.zeros((pr.n, 3))
    np.add.at(full, rows, contrib)
    np.testing.assert_allclose(b, a.bfixed() + full[free], atol=1e-12 * np.abs(b).max())


def test_rigid_motion_is_exact_in_fixed_rest_mode():
    pr = meshgen.sphere_problem(10, 16)
    a = oracle.Arap(pr.xyz, pr.tri)
    idx, _ = pr.constraints()
    a.set_constraints(idx, pr.xyz[idx])
    assert a.deform(1, rest_mode=1)
    np.testing.assert_allclose(a.mesh, pr.xyz, atol=1e-13)  # constraints at rest: p' = p
    Q = meshgen.rot_x(0.7) @ np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    t = np.array([0.3, -1.0, 2.0])
    moved = pr.xyz @ Q.T + t
    a.mesh = moved.copy()
    a.set_constraints(idx, moved[idx])
    assert a.deform(1, rest_mode=1)
    np.testing.assert_allclose(a.mesh, moved, atol=1e-12)
    np.testing.assert_allclose(a.rotations(), np.broadcast_to(Q, (pr.n, 3, 3)), atol=1e-12)
    assert a.energy() < 1e-24


def test_rigid_constraints_converge_to_rigid_motion():
    pr = meshgen.sphere_problem(10, 16)
    a = oracle.Arap(pr.xyz, pr.tri)
    idx, _ = pr.constraints()
    Q = meshgen.rot_x(0.4)
    t = np.array([0.0, 0.2, -0.1])
    a.set_constraints(idx, pr.xyz[idx] @ Q.T + t)
    ok, en = a.deform(300, energy=True)
    assert en[-1] < 1e-8 * en[0]
    np.testing.assert_allclose(a.mesh, pr.xyz @ Q.T + t, atol=1e-5)


def test_deform_semantics():
    pr = meshgen.sphere_problem(8, 12)
    # no constraints: returns true, mesh untouched (arap.h:113-114)
    a = oracle.Arap(pr.xyz, pr.tri)
    assert a.deform(5) is True
    np.testing.assert_array_equal(a.mesh, pr.xyz)
    # deform(0): rest pose with targets pasted (arap.h:122-135)
    idx, tgt = pr.constraints()
    a.set_constraints(idx, tgt)
    assert a.deform(0)
    exp = pr.xyz.copy()
    exp[idx] = tgt
    np.testing.assert_array_equal(a.mesh, exp)
    # every dirty call re-reads the mesh as the new rest pose (arap.h:102-108,162-168)
    a.deform(4)
    after = a.mesh.copy()
    a.set_constraint(idx[0], tgt[0])
    a.deform(0)
    np.testing.assert_array_equal(a.rest(), after)
    # out-of-range index is an error (undefined behaviour in the reference)
    with pytest.raises(IndexError):
        a.set_constraint(pr.n, [0, 0, 0])
    # all vertices constrained: targets written back
    b = oracle.Arap(pr.xyz, pr.tri)
    b.set_constraints(np.arange(pr.n), pr.xyz + 1.0)
    assert b.deform(2)
    np.testing.assert_allclose(b.mesh, pr.xyz + 1.0)


def test_singular_system_reports_false():
    """A component without constraints makes L singular: deform returns false (arap.h:116-117)."""
    x1, t1 = meshgen.uv_sphere(6, 8)
    x2 = x1 + [5.0, 0, 0]
    xyz = np.concatenate([x1, x2])
    tri = np.concatenate([t1, t1 + x1.shape[0]])
    a = oracle.Arap(xyz, tri)
    a.set_constraint(0, xyz[0] + [0, 0, 0.1])
    a.set_constraint(x1.shape[0] - 1, xyz[x1.shape[0] - 1])
    assert a.deform(2) is False
