"""md_mesh_normals_noalloc (the entry point that allocates nothing) against md_mesh_normals,
through the C ABI: bitwise equal on two calls in a row, the second on a smaller mesh (the
workspace keeps the first call's contents), then the first mesh again; meshes beyond the
workspace, and unprepared meshes are refused before any device work."""
import numpy as np
import pytest

import meshgen
import mdarap

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    if mdarap.device_count() == 0:
        pytest.fail("no GPU visible: the HIP path cannot run")
    c = mdarap.Context(0)
    yield c
    c.close()


def _deformed(ctx, pr, iters=3):
    idx, tgt = pr.constraints()
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config(precision=mdarap.PREC_FP32_REFINE))
    m.set_constraints(idx, tgt)
    assert m.deform(iters)
    return m


def test_noalloc_normals_equal_allocating_entry_point(ctx):
    big = meshgen.noisy_icosphere_problem(6, 0.01, 99)     # 40,962 vertices
    small = meshgen.sphere_problem(20, 40)                  # 762 vertices
    say("case: no-alloc normals, n=%d then n=%d" % (big.n, small.n))
    mb = _deformed(ctx, big)
    ms = _deformed(ctx, small)
    nb = mb.normals()
    ns = ms.normals()
    got_b = mb.normals_noalloc()
    got_s = ms.normals_noalloc()          # second call, smaller input
    got_b2 = mb.normals_noalloc()
    assert np.array_equal(got_b, nb)
    assert np.array_equal(got_s, ns)
    assert np.array_equal(got_b2, nb)
    # unit normals of a closed sphere-like surface point outwards
    assert np.allclose(np.linalg.norm(got_s, axis=1), 1.0, atol=1e-12)
    mb.close()
    ms.close()


def test_noalloc_normals_refusals(ctx):
    # beyond the 64 MiB workspace: 24 F + 60 n bytes > 64 MiB (checked from the sizes alone)
    n = 700000
    F = 1400000
    assert 24 * F + 60 * n > 64 * 2**20
    tri = np.stack([np.arange(F) % (n - 2), np.arange(F) % (n - 2) + 1, np.arange(F) % (n - 2) + 2], 1)
    xyz = np.zeros((n, 3))
    m = mdarap.Mesh(ctx, xyz, tri.astype(np.int32), mdarap.config())
    out = np.zeros(3 * n)
    rc = mdarap.lib().md_mesh_normals_noalloc(m.handle, out)
    assert rc == mdarap.MD_ERR_INVALID
    assert b"workspace" in mdarap.lib().md_last_error()
    m.close()
    # a mesh that was never prepared
    pr = meshgen.sphere_problem(6, 8)
    m = mdarap.Mesh(ctx, pr.xyz, pr.tri, mdarap.config())
    rc = mdarap.lib().md_mesh_normals_noalloc(m.handle, np.zeros(3 * pr.n))
    assert rc == mdarap.MD_ERR_STATE
    m.close()
