"""The C++ drop-in (include/deform/arap.h) used the way the reference's callers use it.

tests/cpp/test_dropin.cpp mirrors tests/test_cotan.cpp:20-54 (unit-square cotan weights
through a PrivateAccessor) and examples/deform_sphere.cpp:71-91 (setConstraint + deform
every frame); its printed positions are compared with the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import meshgen
import mdarap
import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
EXE = os.path.join(CPP, "test_dropin")


def build():
    subprocess.check_call(["make", "-s", "-C", CPP])


def run(*args):
    p = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


def test_dropin_compiles_against_headers():
    build()
    assert os.access(EXE, os.X_OK)


def test_dropin_reports_missing_device():
    if mdarap.device_count() > 0:
        pytest.skip("a GPU is present")
    build()
    rc, out, err = run("cotan")
    assert rc == 70 and "no GPU" in err


@pytest.mark.gpu
def test_dropin_cotan_known_answer():
    build()
    rc, out, err = run("cotan")
    assert rc == 0, err
    assert out.strip().endswith("isApprox 1")


@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol", [("sphere", 1e-9), ("sphere32", 1e-5)])
def test_dropin_sphere_matches_oracle(mode, tol):
    build()
    rc, out, err = run(mode, 20, 40, 6)
    assert rc == 0, err
    got = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    pr = meshgen.sphere_problem(20, 40)
    idx, tgt = pr.constraints()
    a = oracle.Arap(pr.xyz, pr.tri)
    a.set_constraints(idx, tgt)
    assert a.deform(6)
    assert meshgen.relative_error(got, a.mesh) < tol


@pytest.mark.gpu
def test_dropin_frames_match_oracle():
    build()
    rc, out, err = run("frames", 12, 24, 5)
    assert rc == 0, err
    got = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    xyz, tri = meshgen.uv_sphere(12, 24)
    a = oracle.Arap(xyz, tri)
    n = xyz.shape[0]
    a.set_constraint(n - 1, xyz[n - 1])
    ph = -np.pi / 2
    for _ in range(5):
        ph += 0.3
        a.set_constraint(0, xyz[0] + [0, 0, np.sin(ph)])
        assert a.deform(5)
    assert meshgen.relative_error(got, a.mesh) < 1e-9
