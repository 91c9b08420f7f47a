"""CPU-side checks of libmdarap.so: it loads, exports every function include/md_arap.h
declares, reports errors without a GPU, and its host analysis (ordering, supernodal
symbolic, FP64 block LDL^T) solves the ARAP system like scipy and the oracle."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import meshgen
import mdarap
import oracle


def header_functions():
    text = open(mdarap.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(md_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = mdarap.lib()
    names = header_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    # the ctypes binding covers exactly the declared functions
    assert sorted(mdarap.exported_names()) == names


def test_version_and_errors_without_device():
    assert b"gfx950" in mdarap.lib().md_version()
    if mdarap.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(mdarap.MDError) as e:
        mdarap.Context(0)
    assert e.value.code == mdarap.MD_ERR_NO_DEVICE


def _problem_system(pr):
    rp, ci, w = oracle.cotan_csr(pr.xyz, pr.tri)
    n = pr.n
    W = sp.csr_matrix((w, ci, rp), shape=(n, n))
    con = np.zeros(n, bool)
    con[pr.anchors] = True
    con[pr.handles] = True
    free = np.nonzero(~con)[0]
    deg = np.asarray(W.sum(axis=1)).ravel()
    L = (sp.diags(deg[free]) - W[free][:, free]).tocsr()
    L.sort_indices()
    return L, pr.xyz[free]


@pytest.mark.parametrize("make", [meshgen.sphere_problem, lambda: meshgen.bar_problem(80, 8, 8),
                                  lambda: meshgen.noisy_icosphere_problem(4)])
@pytest.mark.parametrize("budget", [0, 64, 512, 1024])
@pytest.mark.parametrize("pairs", [False, True])
def test_host_block_ldlt_matches_scipy(make, budget, pairs):
    """The device sweeps (level tiles, and bottom subtrees under each LDS budget), emulated
    on the host with every index bounds-checked and LDS phases race-checked, solve L x = b."""
    pr = make()
    L, fx = _problem_system(pr)
    b = np.random.default_rng(1).standard_normal((3, L.shape[0]))
    x, panel_entries, levels = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b, subtree_budget=budget,
                                                        level_pairs=pairs)
    ref = spla.spsolve(L.tocsc(), b.T).T
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    assert np.abs(L @ x.T - b.T).max() < 1e-10 * np.abs(b).max()
    assert panel_entries >= L.nnz // 2
    assert levels >= 1 if budget == 0 else levels >= 0
    if budget == 0 and not pairs:
        _, _, lv_default = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b)
        assert lv_default < levels  # the default budget takes the bottom levels into subtrees
    if pairs:
        _, _, lv_plain = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b, subtree_budget=budget)
        assert levels <= lv_plain  # paired levels never add launches


@pytest.mark.parametrize("make", [meshgen.sphere_problem, lambda: meshgen.bar_problem(80, 8, 8),
                                  lambda: meshgen.noisy_icosphere_problem(4)])
def test_device_factorization_emulation_matches_scipy(make):
    """The device numeric factorization, emulated on the host with every index checked (front
    assembly, children's update matrices, Cholesky, inverse, G, update matrix), factors L: its
    solves match scipy and the host multifrontal factorization."""
    pr = make()
    L, fx = _problem_system(pr)
    b = np.random.default_rng(3).standard_normal((3, L.shape[0]))
    x_dev, pe, lv = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b, device_factor_emulation=True)
    x_pair, _, _ = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b, device_factor_emulation=True,
                                            level_pairs=True)
    np.testing.assert_allclose(x_pair, spla.spsolve(L.tocsc(), b.T).T, rtol=1e-9,
                               atol=1e-11 * np.abs(x_dev).max())  # wide merged fronts
    x_host, _, _ = mdarap.host_factor_solve(L.indptr, L.indices, L.data, fx, b)
    ref = spla.spsolve(L.tocsc(), b.T).T
    np.testing.assert_allclose(x_dev, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    np.testing.assert_allclose(x_dev, x_host, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_device_factorization_emulation_dense_blocks():
    """A dense SPD matrix of 150 unknowns is one supernode of three column blocks (64, 64, 22):
    the emulated blocked Cholesky, its triangular inverse (diagonal, off-diagonal and partial
    blocks) and D^{-1} give the dense solve."""
    rng = np.random.default_rng(5)
    n = 150
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    S = sp.csr_matrix(A)
    b = rng.standard_normal((2, n))
    xyz = rng.standard_normal((n, 3))
    x, pe, _ = mdarap.host_factor_solve(S.indptr, S.indices, S.data, xyz, b, subtree_budget=0,
                                        device_factor_emulation=True)
    assert pe == n * n
    np.testing.assert_allclose(x, np.linalg.solve(A, b.T).T, rtol=1e-11, atol=1e-13)


def test_host_block_ldlt_dense_small():
    """A dense SPD matrix is one supernode: the block LDL^T is D^{-1} itself."""
    rng = np.random.default_rng(2)
    n = 40
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    S = sp.csr_matrix(A)
    b = rng.standard_normal(n)
    xyz = rng.standard_normal((n, 3))
    for budget, levels in ((0, 1), (-1, 0)):  # one level tile, or one subtree (no level)
        x, pe, lv = mdarap.host_factor_solve(S.indptr, S.indices, S.data, xyz, b, subtree_budget=budget)
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-11)
        assert lv == levels and pe == n * n


def test_host_block_ldlt_rejects_singular():
    x1, t1 = meshgen.uv_sphere(6, 8)
    n = x1.shape[0]
    rp, ci, w = oracle.cotan_csr(x1, t1)
    W = sp.csr_matrix((w, ci, rp), shape=(n, n))
    L = (sp.diags(np.asarray(W.sum(axis=1)).ravel()) - W).tocsr()  # no constraint: singular
    L.sort_indices()
    with pytest.raises(mdarap.MDError) as e:
        mdarap.host_factor_solve(L.indptr, L.indices, L.data, x1, np.ones(n))
    assert e.value.code == mdarap.MD_ERR_FACTOR


# ---- the local step's rotation (kernel code compiled for the host) ------------------------

def test_rotation_matches_svd_golden():
    """The characteristic-polynomial rotation (Newton on Horn's quartic, adjugate eigenvector,
    one Rayleigh-quotient refinement) against numpy.linalg.svd's V diag(1,1,det) U^T on the
    fixture's covariances (reflections and near-rank-2 included): 1e-12 in FP64, 2e-5 in FP32."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "svd_rotations.npz"))
    for fp64, tol in ((True, 1e-12), (False, 2e-5)):
        q, fb = mdarap.host_best_rotation(g["cov"], fp64=fp64)
        R = mdarap.quat_to_matrix(q)
        np.testing.assert_allclose(R, g["R"], atol=tol)
        assert np.all(q[:, 3] >= 0)
        assert fb.sum() <= 2


def test_rotation_planar_rank1_and_zero_covariances():
    rng = np.random.default_rng(12)
    # planar neighbourhoods: rank-2 covariances, a well separated top eigenvalue
    C = []
    for _ in range(200):
        a = rng.standard_normal((6, 3))
        a[:, 2] = 0.0
        Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
        e = a @ Q.T + 1e-3 * rng.standard_normal((6, 3))
        C.append(a.T @ e)
    C = np.array(C)
    ref = np.array([oracle.svd_rotation(c) for c in C])
    q, fb = mdarap.host_best_rotation(C, fp64=True)
    assert not fb.any()
    np.testing.assert_allclose(mdarap.quat_to_matrix(q), ref, atol=1e-12)
    # rank 1 (collinear edges): the top eigenvalue is double, the Jacobi sweeps decide; any
    # rotation taking the edge direction onto the deformed one maximises the trace
    a = rng.standard_normal(3)
    e = rng.standard_normal(3)
    c = np.outer(a, e)
    q, fb = mdarap.host_best_rotation(c[None], fp64=True)
    assert fb[0]
    R = mdarap.quat_to_matrix(q)[0]
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
    assert np.trace(R @ c) == pytest.approx(np.linalg.norm(a) * np.linalg.norm(e), rel=1e-10)
    # zero covariance: the previous rotation stays
    prev = np.array([[0.1, -0.2, 0.3, 0.9]])
    prev /= np.linalg.norm(prev)
    q, _ = mdarap.host_best_rotation(np.zeros((1, 3, 3)), prev=prev, fp64=True)
    np.testing.assert_allclose(q, prev, atol=1e-15)
