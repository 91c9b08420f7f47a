"""GPU parity of the HIP path against the FP64 oracle, through the C ABI.

Tolerances (position error = max |gpu - oracle| / max |oracle| over all coordinates):
  * FP64 path: 1e-9 (positions), 1e-12 relative (cotan weights), 1e-9 (rotations, energy).
  * FP32 local step + FP32 block LDL^T + FP64 residual refinement: 1e-5 (positions).
"""
import os
import sys

import numpy as np
import pytest

import meshgen
import mdarap
import oracle

pytestmark = pytest.mark.gpu

POS_TOL_FP64 = 1e-9
POS_TOL_FP32 = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    if mdarap.device_count() == 0:
        pytest.fail("no GPU visible: the HIP path cannot run")
    c = mdarap.Context(0)
    yield c
    c.close()


def gold(golden_dir, name):
    return np.load(os.path.join(golden_dir, name + ".npz"))


def run_gpu(ctx, xyz, tri, idx, tgt, iters, precision=mdarap.PREC_FP64, **kw):
    m = mdarap.Mesh(ctx, xyz, tri, mdarap.config(precision=precision, **kw))
    m.set_constraints(idx, tgt)
    ok = m.deform(iters)
    return m, ok


# ---- cotan weights (arap.h:182-239) -------------------------------------------------------

def test_cotan_known_answer(ctx, golden_dir):
    """tests/test_cotan.cpp:20-54 through the PrivateAccessor-equivalent export."""
    say("case: unit square cotan")
    g = gold(golden_dir, "unit_square")
    m = mdarap.Mesh(ctx, g["xyz"], g["tri"])
    assert m.deform(0) is True  # no constraints: returns true (arap.h:113-114)
    rp, ci, v = m.cotan()
    import scipy.sparse as sp
    W = sp.csr_matrix((v, ci, rp), shape=(4, 4)).toarray().astype(np.float32)
    exp = g["expected"]
    assert np.linalg.norm(W - exp) <= 1e-4 * min(np.linalg.norm(W), np.linalg.norm(exp))
    np.testing.assert_array_equal(m.mesh, g["xyz"])  # mesh untouched


@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "plane_obj", "uv_sphere"])
def test_cotan_matches_oracle(ctx, golden_dir, name):
    say("case: cotan", name)
    g = gold(golden_dir, name)
    m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], 0)
    assert ok
    rp, ci, v = m.cotan()
    assert np.array_equal(rp, g["rp"]) and np.array_equal(ci, g["ci"])
    np.testing.assert_allclose(v, g["w"], rtol=1e-12, atol=0)


def _degenerate_grid():
    """A 7 x 7 grid patch with degenerate faces: two coincident vertices (zero-length edge,
    zero-area faces), a needle (a vertex moved almost onto an edge), a NaN-prone flat triangle,
    and obtuse corners; anchors on the first row, handles on the last row lifted 0.3."""
    k = 7
    g = np.arange(k, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="xy")
    xyz = np.stack([X.ravel(), Y.ravel(), 0.01 * np.sin(X.ravel() * 1.3 + Y.ravel())], axis=1) / (k - 1)
    tri = []
    for j in range(k - 1):
        for i in range(k - 1):
            a, b, c, d = j * k + i, j * k + i + 1, (j + 1) * k + i, (j + 1) * k + i + 1
            tri += [(a, b, d), (a, d, c)]
    tri = np.array(tri, np.int32)
    xyz[17] = xyz[18]                                  # coincident vertices
    xyz[31] = 0.5 * (xyz[30] + xyz[32]) + [0, 1e-9, 0]  # needle: almost on the edge 30-32
    xyz[38, 1] += 0.4 / (k - 1)                         # obtuse corners around vertex 38
    anchors = np.arange(k)
    handles = np.arange(k * (k - 1), k * k)
    idx = np.concatenate([anchors, handles]).astype(np.int32)
    tgt = xyz[idx].copy()
    tgt[len(anchors):, 2] += 0.3
    return xyz, tri, idx, tgt


def test_cotan_and_deform_on_degenerate_faces(ctx):
    """Degenerate geometry (arap.h:199-218 clamps: max(1e-8, .) lengths and Heron area,
    max(1e-10, .) cotangents): the device weights equal the oracle's, and the FP64 deformation
    follows the oracle (loosely: the clamped weights make the system ill-conditioned)."""
    say("case: degenerate faces")
    xyz, tri, idx, tgt = _degenerate_grid()
    rp0, ci0, w0 = oracle.cotan_csr(xyz, tri)
    assert np.all(w0 > 0) and np.any(w0 < 1e-6)  # the clamps are exercised
    m, ok = run_gpu(ctx, xyz, tri, idx, tgt, 5)
    assert ok
    rp, ci, v = m.cotan()
    assert np.array_equal(rp, rp0) and np.array_equal(ci, ci0)
    np.testing.assert_allclose(v, w0, rtol=1e-12, atol=0)
    a = oracle.Arap(xyz, tri)
    a.set_constraints(idx, tgt)
    assert a.deform(5)
    assert np.all(np.isfinite(m.mesh))
    assert meshgen.relative_error(m.mesh, a.mesh) < 1e-6


@pytest.mark.parametrize("budget", [0, 256])
def test_disconnected_components_match_oracle(ctx, budget):
    """Three disconnected components in one mesh (a UV sphere, a small bar, a small icosphere),
    each with its own anchors and handles: the factor is a forest with several roots (roots
    without off rows, whose backward tiles gather their right-hand side); FP64 at 1e-9 and
    FP32 + refinement at 1e-5 against the oracle."""
    say("case: disconnected components, budget", budget)
    parts = [meshgen.sphere_problem(12, 24), meshgen.bar_problem(40, 4, 4), meshgen.noisy_icosphere_problem(3, 0.01, 7)]
    xyz, tri, idx, tgt, base = [], [], [], [], 0
    for k, pr in enumerate(parts):
        shift = np.array([3.0 * k, 0.0, 0.0])
        xyz.append(pr.xyz + shift)
        tri.append(pr.tri + base)
        ci, ct = pr.constraints()
        idx.append(ci + base)
        tgt.append(ct + shift)
        base += pr.n
    xyz, tri = np.concatenate(xyz), np.concatenate(tri).astype(np.int32)
    idx, tgt = np.concatenate(idx).astype(np.int32), np.concatenate(tgt)
    a = oracle.Arap(xyz, tri)
    a.set_constraints(idx, tgt)
    assert a.deform(8)
    for prec, tol in ((mdarap.PREC_FP64, POS_TOL_FP64), (mdarap.PREC_FP32_REFINE, POS_TOL_FP32)):
        m, ok = run_gpu(ctx, xyz, tri, idx, tgt, 8, precision=prec, subtree_budget=budget)
        assert ok
        err = meshgen.relative_error(m.mesh, a.mesh)
        say("  precision", prec, "levels", m.stats()["levels"], "rel err", err)
        assert err < tol


# ---- local step (arap.h:354-384) ----------------------------------------------------------

@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "plane_obj", "uv_sphere"])
def test_first_rotations_match_oracle(ctx, golden_dir, name):
    say("case: rotations", name)
    g = gold(golden_dir, name)
    m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], 1)
    R = m.rotations()
    np.testing.assert_allclose(R, g["R1"], atol=1e-9)
    m32, _ = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], 1, precision=mdarap.PREC_FP32_REFINE)
    np.testing.assert_allclose(m32.rotations(), g["R1"], atol=2e-4)


# ---- whole deform (arap.h:101-138) --------------------------------------------------------

@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "plane_obj", "uv_sphere"])
def test_deform_fp64_matches_golden(ctx, golden_dir, name):
    say("case: deform fp64", name)
    g = gold(golden_dir, name)
    m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], int(g["iters"]))
    assert ok
    err = meshgen.relative_error(m.mesh, g["pos"])
    say("  rel err", err)
    assert err < POS_TOL_FP64
    # energy of the final state (rotations of the last local step)
    assert m.energy() == pytest.approx(float(g["energy"][-1]), rel=1e-9)


@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "plane_obj", "uv_sphere"])
def test_deform_fp32_refine_matches_golden(ctx, golden_dir, name):
    say("case: deform fp32+refine", name)
    g = gold(golden_dir, name)
    m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], int(g["iters"]),
                    precision=mdarap.PREC_FP32_REFINE)
    assert ok
    err = meshgen.relative_error(m.mesh, g["pos"])
    say("  rel err", err)
    assert err < POS_TOL_FP32


@pytest.mark.parametrize("name", ["sphere_obj", "bar_obj", "uv_sphere"])
@pytest.mark.parametrize("budget", [0, 64, 1024])
def test_deform_sweep_layouts_match_golden(ctx, golden_dir, name, budget):
    """Every sweep layout (level tiles only, small and large bottom subtrees swept in LDS)
    gives the oracle's positions: 1e-9 in FP64, 1e-5 in FP32 + refinement."""
    say("case: sweep layout", name, "budget", budget)
    g = gold(golden_dir, name)
    for prec, tol in ((mdarap.PREC_FP64, POS_TOL_FP64), (mdarap.PREC_FP32_REFINE, POS_TOL_FP32)):
        m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], int(g["iters"]), precision=prec,
                        subtree_budget=budget)
        assert ok
        err = meshgen.relative_error(m.mesh, g["pos"])
        say("  precision", prec, "levels", m.stats()["levels"], "rel err", err)
        assert err < tol


@pytest.mark.parametrize("which", ["sphere", "ico", "bar", "bar16k", "bar67k", "ico164k"])
def test_device_factorization_matches_host(ctx, which):
    """The device numeric factorization (FP64, multifrontal on the GPU) and the host one give
    the same iterates: FP64 positions after 3 iterations agree to 1e-11 relative, and both meet
    the oracle at 1e-9. Every case has pivot blocks past the first
    (factored inside the trailing update's launch); the three larger ones also split the top
    heights' products over K (2 splits on the 16.8k bar; 2 and 4 on the 67k bar; 2 to 6 on the
    164k icosphere)."""
    pr = {"sphere": lambda: meshgen.sphere_problem(), "ico": lambda: meshgen.noisy_icosphere_problem(5),
          "bar": lambda: meshgen.bar_problem(120, 12, 12), "bar16k": lambda: meshgen.bar_problem(200, 20, 20),
          "bar67k": lambda: meshgen.bar_problem(400, 40, 40),
          "ico164k": lambda: meshgen.noisy_icosphere_problem(7)}[which]()
    say("case: device factorization", pr.name, "n", pr.n)
    idx, tgt = pr.constraints()
    out = []
    for dev in (True, False):
        m, ok = run_gpu(ctx, pr.xyz, pr.tri, idx, tgt, 3, factor_device=dev)
        assert ok
        say("  factor_device", dev, "factor ms %.1f" % m.stats()["factor_ms"])
        out.append(m.mesh)
    assert meshgen.relative_error(out[0], out[1]) < 1e-11
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    assert a.deform(3)
    assert meshgen.relative_error(out[0], a.mesh) < POS_TOL_FP64


def test_device_factorization_bitwise_repeatable(ctx):
    """Two fresh meshes with the same frame give bitwise-identical positions after 3 iterations:
    the device factorization has no floating-point atomics, and the partial products of the top
    heights, split over K (2 and 4 ways on this 67k bar), are summed in split order."""
    pr = meshgen.bar_problem(400, 40, 40)
    say("case: repeatable device factorization", pr.name, "n", pr.n)
    idx, tgt = pr.constraints()
    outs = []
    for _ in range(2):
        m, ok = run_gpu(ctx, pr.xyz, pr.tri, idx, tgt, 3)
        assert ok
        outs.append(np.array(m.mesh, copy=True))
        m.close()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_energy_trace_monotone(ctx):
    say("case: energy trace")
    pr = meshgen.sphere_problem()
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri)
    m.set_constraints(idx, tgt)
    m.deform(0)
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    _, en = a.deform(8, energy=True)
    got = []
    for _ in range(8):
        m.iterate(1)
        got.append(m.energy())
    np.testing.assert_allclose(got, en, rtol=1e-9)
    assert np.all(np.diff(got) <= 1e-12 * got[0])


# ---- deform() semantics --------------------------------------------------------------------

def test_semantics(ctx):
    say("case: semantics")
    pr = meshgen.sphere_problem(8, 12)
    idx, tgt = pr.constraints()
    # deform(0) pastes targets into the rest pose (arap.h:122-135)
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri)
    m.set_constraints(idx, tgt)
    assert m.deform(0)
    exp = pr.xyz.copy()
    exp[idx] = tgt
    np.testing.assert_array_equal(m.mesh, exp)
    # out-of-range index is an error
    with pytest.raises(mdarap.MDError) as e:
        m.set_constraint(pr.n, [0, 0, 0])
    assert e.value.code == mdarap.MD_ERR_INVALID
    # all vertices constrained: targets are written back
    m2 = mdarap.Mesh(ctx, pr.xyz, pr.tri)
    m2.set_constraints(np.arange(pr.n), pr.xyz + 1.0)
    assert m2.deform(3)
    np.testing.assert_allclose(m2.mesh, pr.xyz + 1.0)


def test_singular_returns_false(ctx):
    say("case: singular")
    x1, t1 = meshgen.uv_sphere(6, 8)
    xyz = np.concatenate([x1, x1 + [5.0, 0, 0]])
    tri = np.concatenate([t1, t1 + x1.shape[0]])
    m = mdarap.Mesh(ctx, xyz, tri)
    m.set_constraint(0, xyz[0] + [0, 0, 0.1])
    m.set_constraint(x1.shape[0] - 1, xyz[x1.shape[0] - 1])
    before = m.mesh.copy()
    assert m.deform(2) is False
    np.testing.assert_array_equal(m.mesh, before)


@pytest.mark.parametrize("rest_mode", [mdarap.REST_REFERENCE, mdarap.REST_FIXED])
def test_interactive_frames_match_oracle(ctx, golden_dir, rest_mode):
    """Frame loop of examples/deform_sphere.cpp:71-91: setConstraint + deform(5) per frame."""
    say("case: frames rest_mode", rest_mode)
    g = gold(golden_dir, "sphere_obj")
    xyz, tri = g["xyz"], g["tri"]
    m = mdarap.Mesh(ctx, xyz, tri, mdarap.config(rest_mode=rest_mode))
    a = oracle.Arap(xyz, tri)
    m.set_constraint(37, xyz[37])
    a.set_constraint(37, xyz[37])
    for frame in range(6):
        tgt = xyz[32] + [0, 0, np.sin(-np.pi / 2 + 0.3 * frame)]
        m.set_constraint(32, tgt)
        a.set_constraint(32, tgt)
        assert m.deform(5)
        assert a.deform(5, rest_mode=rest_mode)
        err = meshgen.relative_error(m.mesh, a.mesh)
        assert err < POS_TOL_FP64, (frame, err)


def test_reference_mode_keeps_factor_workspace(ctx, golden_dir):
    """A mesh in reference rest mode refactors on every dirty deform() and keeps the FP64
    factorization workspace between calls (md_mesh_stats.device_bytes counts it); a fixed-rest
    mesh frees it after its factorization. Frames that reuse the workspace match the oracle."""
    say("case: factorization workspace by rest mode")
    g = gold(golden_dir, "bar_obj")
    xyz, tri, idx, tgt = g["xyz"], g["tri"], g["con_idx"], g["con_xyz"]
    ref, ok = run_gpu(ctx, xyz, tri, idx, tgt, 2, rest_mode=mdarap.REST_REFERENCE)
    assert ok
    fix, ok = run_gpu(ctx, xyz, tri, idx, tgt, 2, rest_mode=mdarap.REST_FIXED)
    assert ok
    extra = ref.stats()["device_bytes"] - fix.stats()["device_bytes"]
    assert extra > 0, extra
    a = oracle.Arap(xyz, tri)
    a.set_constraints(idx, tgt)
    assert a.deform(2)
    for frame in range(3):
        t = tgt + 0.02 * (frame + 1)
        ref.set_constraints(idx, t)
        a.set_constraints(idx, t)
        assert ref.deform(3)
        assert a.deform(3)
        assert ref.stats()["device_bytes"] - fix.stats()["device_bytes"] == extra
        err = meshgen.relative_error(ref.mesh, a.mesh)
        assert err < POS_TOL_FP64, (frame, err)


def test_large_mesh_staged_position_copies_match_oracle(ctx):
    """The 164k icosphere (3.9 MB of positions): from a mesh's second download on, its positions
    come back through the pinned, chunked staging. Reference rest mode over three frames, then
    fixed rest mode with the constrained set changed (the rest pose and the current positions
    uploaded in one preparation): every frame matches the oracle."""
    pr = meshgen.noisy_icosphere_problem(7)
    idx, tgt = pr.constraints()
    say("case: staged position downloads, reference frames", pr.name, "n", pr.n)
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(rest_mode=mdarap.REST_REFERENCE))
    a = oracle.Arap(pr.xyz, pr.tri)
    for frame in range(3):
        t = tgt + 0.01 * frame
        m.set_constraints(idx, t)
        a.set_constraints(idx, t)
        assert m.deform(2)
        assert a.deform(2)
        err = meshgen.relative_error(m.mesh, a.mesh)
        say("  frame", frame, "rel err", err)
        assert err < POS_TOL_FP64, (frame, err)
    m.close()
    say("case: staged position downloads, fixed rest with a new constrained set")
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(rest_mode=mdarap.REST_FIXED))
    a = oracle.Arap(pr.xyz, pr.tri)
    extra = int(np.setdiff1d(np.arange(pr.n), idx)[0])
    for obj in (m, a):
        obj.set_constraints(idx, tgt)
    assert m.deform(2)
    assert a.deform(2, rest_mode=1)
    for obj in (m, a):
        obj.set_constraint(extra, pr.xyz[extra] + [0.0, 0.0, 0.05])
    assert m.deform(2)
    assert a.deform(2, rest_mode=1)
    err = meshgen.relative_error(m.mesh, a.mesh)
    say("  rel err", err)
    assert err < POS_TOL_FP64
    m.close()


def test_fixed_rest_new_constraint_set_matches_oracle(ctx, golden_dir):
    """Fixed rest mode with the constrained set changed between frames: the rest pose and its
    weights stay those of the first call while the current mesh is the initial guess (the
    library uploads the kept rest pose and the caller's positions separately in that case)."""
    say("case: fixed rest, constrained set changed after the first frame")
    g = gold(golden_dir, "sphere_obj")
    xyz, tri = g["xyz"], g["tri"]
    m = mdarap.Mesh(ctx, xyz, tri, mdarap.config(rest_mode=mdarap.REST_FIXED))
    a = oracle.Arap(xyz, tri)
    for obj in (m, a):
        obj.set_constraint(37, xyz[37])
        obj.set_constraint(32, xyz[32] + [0.0, 0.0, 0.3])
    assert m.deform(5)
    assert a.deform(5, rest_mode=1)
    for frame, v in enumerate((10, 200, 400)):
        for obj in (m, a):
            obj.set_constraint(v, xyz[v] + [0.05 * (frame + 1), 0.0, 0.0])
        assert m.deform(5)
        assert a.deform(5, rest_mode=1)
        err = meshgen.relative_error(m.mesh, a.mesh)
        assert err < POS_TOL_FP64, (frame, err)


# ---- full-size workloads ------------------------------------------------------------------

@pytest.mark.parametrize("which,iters", [("bar", 50)])
def test_large_workloads_fp64_and_fp32(ctx, which, iters):
    """The 67k-vertex bar (40 x 40 x 400 quads, one end anchored, the other rotated 90 degrees
    about the long axis and lifted 0.3) over 50 iterations, FP64 and FP32 + refinement against
    the oracle (the 164k icosphere over 100 iterations is in test_gpu_configs.py)."""
    pr = {"bar": meshgen.bar_problem}[which]()
    say("case: full size", pr.name, "n", pr.n)
    idx, tgt = pr.constraints()
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    a.deform(iters)
    for prec, tol in ((mdarap.PREC_FP64, POS_TOL_FP64), (mdarap.PREC_FP32_REFINE, POS_TOL_FP32)):
        m, ok = run_gpu(ctx, pr.xyz, pr.tri, idx, tgt, iters, precision=prec)
        assert ok
        err = meshgen.relative_error(m.mesh, a.mesh)
        say("  precision", prec, "rel err", err, "stats", m.stats())
        assert err < tol


def test_batch_equals_individual(ctx):
    say("case: batch")
    probs = meshgen.mixed_batch(6, 500, 6000, seed=7)
    outs, meshes = [], []
    for pr in probs:
        idx, tgt = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
        m.set_constraints(idx, tgt)
        m.prepare(pr.xyz)
        m.iterate(4)
        outs.append(m.download())
        meshes.append(m)
    bm = []
    for pr in probs:
        idx, tgt = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
        m.set_constraints(idx, tgt)
        m.prepare(pr.xyz)
        bm.append(m)
    b = mdarap.Batch(ctx, bm)
    b.iterate(4)
    for m, ref in zip(bm, outs):
        np.testing.assert_array_equal(m.download(), ref)


def test_graph_replays_equal_direct_launches(ctx):
    """Long runs replay a graph of 8 iterations, the rest a graph of one: any split of the
    iterations gives bitwise what direct launches give, for one mesh and for a batch."""
    say("case: graph replays")
    probs = [meshgen.noisy_icosphere_problem(4)] + meshgen.mixed_batch(3, 500, 4000, seed=11)
    ref = []
    for pr in probs:
        idx, tgt = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, use_graph=False))
        m.set_constraints(idx, tgt)
        m.prepare(pr.xyz)
        m.iterate(22)
        ref.append(m.download())
        m.close()
    pr = probs[0]
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
    m.set_constraints(idx, tgt)
    m.prepare(pr.xyz)
    for k in (3, 19):  # one-iteration graph, then two 8-iteration replays and three single ones
        m.iterate(k)
    np.testing.assert_array_equal(m.download(), ref[0])
    m.close()
    bm = []
    for pr in probs:
        idx, tgt = pr.constraints()
        mm = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
        mm.set_constraints(idx, tgt)
        mm.prepare(pr.xyz)
        bm.append(mm)
    b = mdarap.Batch(ctx, bm)
    b.iterate(16)
    b.iterate(6)
    for mm, r in zip(bm, ref):
        np.testing.assert_array_equal(mm.download(), r)
    b.close()
    for mm in bm:
        mm.close()


def _prepared(ctx, probs, budget=None):
    out = []
    for pr in probs:
        idx, tgt = pr.constraints()
        m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, subtree_budget=budget))
        m.set_constraints(idx, tgt)
        m.prepare(pr.xyz)
        out.append(m)
    return out


def test_mesh_autotune_keeps_state_and_results(ctx):
    """The autotuner leaves the mesh state as it found it and keeps a candidate layout; the
    iterations after it give what a mesh created with that layout gives, bitwise."""
    say("case: mesh autotune")
    pr = meshgen.noisy_icosphere_problem(5)
    (m,) = _prepared(ctx, [pr])
    before = m.download()
    best, ms = m.autotune((0, 128, 1024), iters=2)
    say("  best budget", best, "ms/iter", ms)
    assert best in (0, 128, 1024) and ms > 0
    np.testing.assert_array_equal(m.download(), before)
    m.iterate(5)
    (ref,) = _prepared(ctx, [pr], budget=best)
    ref.iterate(5)
    np.testing.assert_array_equal(m.download(), ref.download())
    # tuning between iterations: the warm-start rotations are restored too
    m2, ref2 = _prepared(ctx, [pr, pr], budget=best)
    m2.iterate(2)
    ref2.iterate(2)
    m2.autotune((0, 1024, best), iters=2)
    m2.autotune((best,), iters=1)
    m2.iterate(3)
    ref2.iterate(3)
    np.testing.assert_array_equal(m2.download(), ref2.download())


def test_factor_autotune_equals_fresh_mesh(ctx):
    """The factor-layout autotuner leaves the freshly prepared state of the chosen (ratio,
    budget); iterations then match a mesh created with that configuration bitwise, and the
    oracle within the FP32 + refinement tolerance."""
    say("case: factor autotune")
    pr = meshgen.noisy_icosphere_problem(4)
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
    m.set_constraints(idx, tgt)
    m.prepare(pr.xyz)
    fresh = m.download()
    ratio, budget, ms = m.autotune_factor((0.3, 1.0), (0, 256), iters=2)
    say("  best ratio", ratio, "budget", budget, "ms/iter", ms)
    assert ratio in (0.3, 1.0) and budget in (0, 256) and ms > 0
    st = m.stats()
    assert st["subtree_budget"] == budget
    np.testing.assert_array_equal(m.download(), fresh)
    with pytest.raises(mdarap.MDError) as e:  # only before the first iteration
        m.iterate(1)
        m.autotune_factor((0.3,), (0,), iters=1)
    assert e.value.code == mdarap.MD_ERR_STATE
    ref = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE, amalg_ratio=ratio,
                                                          subtree_budget=budget))
    ref.set_constraints(idx, tgt)
    ref.prepare(pr.xyz)
    ref.iterate(1)
    m.iterate(4)
    ref.iterate(4)
    got = m.download()
    np.testing.assert_array_equal(got, ref.download())
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    assert a.deform(5)
    assert meshgen.relative_error(got, a.mesh) < POS_TOL_FP32


def test_batch_autotune_and_stale_batch(ctx):
    """Batch autotune: one layout for all meshes, results bitwise those of meshes created with
    it; a batch whose mesh was re-laid out afterwards refuses to launch."""
    say("case: batch autotune")
    probs = meshgen.mixed_batch(5, 500, 8000, seed=13)
    meshes = _prepared(ctx, probs)
    b = mdarap.Batch(ctx, meshes)
    best, ms = b.autotune((0, 512, 1024), iters=2)
    say("  best budget", best, "ms/iter", ms)
    assert best in (0, 512, 1024)
    b.iterate(4)
    refs = _prepared(ctx, probs, budget=best)
    for r in refs:
        r.iterate(4)
    for m, r in zip(meshes, refs):
        np.testing.assert_array_equal(m.download(), r.download())
    meshes[0].autotune((0,), iters=1)
    with pytest.raises(mdarap.MDError) as e:
        b.iterate(1)
    assert e.value.code == mdarap.MD_ERR_STATE
    b.close()


# ---- Jacobi-PCG alternative ------------------------------------------------------------------

@pytest.mark.parametrize("precision,tol", [(mdarap.PREC_FP64, 1e-7), (mdarap.PREC_FP32_REFINE, 1e-5)])
def test_pcg_matches_golden(ctx, golden_dir, precision, tol):
    say("case: pcg", precision)
    g = gold(golden_dir, "uv_sphere")
    m, ok = run_gpu(ctx, g["xyz"], g["tri"], g["con_idx"], g["con_xyz"], int(g["iters"]), precision=precision,
                    solver=mdarap.SOLVER_PCG, pcg_iters=400, pcg_tol=1e-12 if precision == mdarap.PREC_FP64 else 1e-7,
                    refine_steps=2)
    assert ok
    err = meshgen.relative_error(m.mesh, g["pos"])
    say("  rel err", err)
    assert err < tol


def test_vertex_normals(ctx):
    """Normals of the deformed mesh: unit, radial on an undeformed sphere, and equal to the
    normalised sum of incident area-weighted face normals computed independently."""
    say("case: normals")
    xyz, tri = meshgen.icosphere(3)
    m = mdarap.Mesh(ctx, xyz, tri)
    m.set_constraint(0, xyz[0])
    assert m.deform(1)  # one constraint at rest: the sphere stays put
    nrm = m.normals()
    np.testing.assert_allclose(np.linalg.norm(nrm, axis=1), 1.0, rtol=1e-12)
    assert np.min(np.einsum("ij,ij->i", nrm, xyz / np.linalg.norm(xyz, axis=1, keepdims=True))) > 0.995
    pos = m.mesh
    fn = np.cross(pos[tri[:, 1]] - pos[tri[:, 0]], pos[tri[:, 2]] - pos[tri[:, 0]])
    acc = np.zeros_like(pos)
    for k in range(3):
        np.add.at(acc, tri[:, k], fn)
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    np.testing.assert_allclose(nrm, acc, atol=1e-12)


def test_context_outlives_its_meshes_regression():
    """A context destroyed before its meshes, batches or servers stays valid until the last
    of them is destroyed (each holds a reference)."""
    say("case: context lifetime")
    if mdarap.device_count() == 0:
        pytest.fail("no GPU visible")
    c = mdarap.Context(0)
    pr = meshgen.sphere_problem(6, 8)
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(c, pr.xyz, pr.tri, mdarap.config(rest_mode=mdarap.REST_FIXED))
    m.set_constraints(idx, tgt)
    m.prepare()
    m2 = mdarap.Mesh(c, pr.xyz, pr.tri, mdarap.config(rest_mode=mdarap.REST_FIXED))
    m2.set_constraints(idx, tgt)
    m2.prepare()
    b = mdarap.Batch(c, [m])
    s = mdarap.Server(c)
    mid = s.add_mesh(m2)   # a served mesh belongs to the server until it is destroyed
    c.close()          # the caller's handle goes first
    b.iterate(1)       # the objects still work
    t = s.submit(mid, idx[:1], tgt[:1], 1)
    out, _ = s.wait(t, pr.n)
    assert np.all(np.isfinite(out))
    s.close()
    b.close()
    m.iterate(1)
    m2.iterate(1)
    assert np.all(np.isfinite(m.download()))
    assert np.all(np.isfinite(m2.download()))
    m.close()
    m2.close()
