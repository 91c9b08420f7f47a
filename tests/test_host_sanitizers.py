"""The host analysis under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/test_host_asan is md_analysis.cpp and md_host.cpp (ordering, level pairs, one-ring
topology, symbolic factor, device layout for every subtree budget, FP64 factorization on the
host and by the emulation of the device kernels, bounds-checked emulation of the device
sweeps) built with -fsanitize=address,undefined -fno-sanitize-recover=all: any out-of-bounds
access, use after free, leak, signed overflow or misaligned access aborts it. It runs on meshes
with constrained vertices (the analysis of a dirty deform()) and on cotan systems through
md_host_factor_solve, and the solutions must meet the equations (relative residual 1e-9).
"""
import os
import subprocess

import numpy as np
import pytest

import meshgen
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


@pytest.fixture(scope="module")
def prog():
    subprocess.check_call(["make", "-s", "-C", CPP, "test_host_asan"])
    return os.path.join(CPP, "test_host_asan")


def run(prog, mode, path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([prog, mode, path], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, (out.stdout + out.stderr)[-4000:]
    return float(out.stdout.split()[-1])


def write_mesh(path, pr):
    idx, _ = pr.constraints()
    with open(path, "wb") as f:
        np.array([pr.n, pr.F, len(idx)], np.int32).tofile(f)
        pr.tri.astype(np.int32).tofile(f)
        pr.xyz.astype(np.float64).tofile(f)
        np.asarray(idx, np.int32).tofile(f)


def three_components():
    parts = [meshgen.sphere_problem(8, 16), meshgen.bar_problem(20, 3, 3), meshgen.noisy_icosphere_problem(2, 0.01, 3)]
    xyz, tri, anc, han, base = [], [], [], [], 0
    for k, pr in enumerate(parts):
        xyz.append(pr.xyz + [3.0 * k, 0, 0])
        tri.append(pr.tri + base)
        anc.append(pr.anchors + base)
        han.append(pr.handles + base)
        base += pr.n
    xyz = np.concatenate(xyz)
    han = np.concatenate(han)
    return meshgen.Problem("three_components", xyz, np.concatenate(tri), np.concatenate(anc), han, xyz[han])


@pytest.mark.parametrize("which", ["uv_sphere", "bar", "icosphere", "three_components"])
def test_mesh_analysis_under_sanitizers(prog, tmp_path, which):
    pr = {"uv_sphere": lambda: meshgen.sphere_problem(20, 40), "bar": lambda: meshgen.bar_problem(60, 6, 6),
          "icosphere": lambda: meshgen.noisy_icosphere_problem(4, 0.01, 5), "three_components": three_components}[which]()
    path = str(tmp_path / "mesh.bin")
    write_mesh(path, pr)
    res = run(prog, "mesh", path)
    assert res < 1e-9


def test_cotan_system_under_sanitizers(prog, tmp_path):
    pr = meshgen.sphere_problem(16, 32)
    rp, ci, w = oracle.cotan_csr(pr.xyz, pr.tri)
    n = pr.n
    idx, _ = pr.constraints()
    free = np.setdiff1d(np.arange(n), idx)
    pos = -np.ones(n, np.int64)
    pos[free] = np.arange(free.size)
    rows, cols, vals = [], [], []
    for i in free:
        d = 0.0
        for e in range(rp[i], rp[i + 1]):
            j = ci[e]
            d += w[e]
            if pos[j] >= 0:
                rows.append(pos[i]); cols.append(pos[j]); vals.append(-w[e])
        rows.append(pos[i]); cols.append(pos[i]); vals.append(d)
    order = np.lexsort((cols, rows))
    rows, cols, vals = np.asarray(rows)[order], np.asarray(cols, np.int32)[order], np.asarray(vals)[order]
    m = free.size
    rowptr = np.zeros(m + 1, np.int32)
    np.add.at(rowptr, np.asarray(rows) + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    rng = np.random.default_rng(3)
    b = rng.standard_normal((4, m))
    path = str(tmp_path / "system.bin")
    with open(path, "wb") as f:
        np.array([m, len(vals), b.shape[0]], np.int32).tofile(f)
        rowptr.tofile(f)
        cols.tofile(f)
        vals.astype(np.float64).tofile(f)
        pr.xyz[free].astype(np.float64).tofile(f)
        b.tofile(f)
    res = run(prog, "system", path)
    assert res < 1e-9
