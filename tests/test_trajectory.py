"""Pins the host-side handle motion (trajectory.h, deformation_util.h restated).

* tests/test_trajectory.cpp:17-39: a 4-keyframe trajectory starts at pose 0 and ends at
  pose 3 (Eigen isApprox, 1e-3).
* SE3 exp/log are inverse to each other and match scipy's rotation-vector conversions.
* With 4 keyframes the spline is a single cubic Bezier segment: the B-spline basis equals the
  Bernstein polynomials, and the curve passes through every log-pose at its chord-length
  parameter.
* DeformationUtil with origin O: targets = O T O^-1 p0 (deformation_util.h:51-55).
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import trajectory as tj


def is_approx(a, b, prec):
    # Eigen isApprox: ||a - b|| <= prec * min(||a||, ||b||)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def translation(x, y, z):
    T = np.eye(4)
    T[:3, 3] = [x, y, z]
    return T


def rotation_x(a):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec([a, 0, 0]).as_matrix()
    return T


def test_reference_endpoints():
    """tests/test_trajectory.cpp:17-39 with the same four poses."""
    path = tj.TrajectorySE3()
    pose0 = np.eye(4)
    path.add_key_pose(pose0)
    pose1 = translation(0, 0, 1) @ pose0
    path.add_key_pose(pose1)
    pose2 = translation(0, 0, 1) @ rotation_x(math.pi / 4) @ pose1
    path.add_key_pose(pose2)
    pose3 = translation(0, 0, 1) @ pose2
    path.add_key_pose(pose3)
    assert is_approx(path(0.0), pose0, 1e-3)
    assert is_approx(path(1.0), pose3, 1e-3)
    # and much tighter in double precision
    np.testing.assert_allclose(path(1.0), pose3, atol=1e-12)


def test_se3_exp_log_roundtrip_and_scipy():
    rng = np.random.default_rng(4)
    for _ in range(50):
        w = rng.uniform(-1, 1, 3)
        w *= rng.uniform(0, 3.0) / np.linalg.norm(w)  # principal branch: angle < pi
        ups = rng.standard_normal(3)
        T = tj.se3_exp(np.concatenate([ups, w]))
        np.testing.assert_allclose(T[:3, :3], Rotation.from_rotvec(w).as_matrix(), atol=1e-12)
        xi = tj.se3_log(T)
        np.testing.assert_allclose(xi, np.concatenate([ups, w]), atol=1e-9)
    # pure translation: V = I
    T = tj.se3_exp([1.0, 2.0, 3.0, 0, 0, 0])
    np.testing.assert_allclose(T[:3, 3], [1, 2, 3])


def test_four_keyframes_are_one_bezier_segment():
    u = np.array([0.0, 0.3, 0.55, 1.0])
    knots = tj.knot_averaging(u, 3)
    np.testing.assert_array_equal(knots, [0, 0, 0, 0, 1, 1, 1, 1])
    for t in np.linspace(0, 1, 11):
        bern = [math.comb(3, j) * t ** j * (1 - t) ** (3 - j) for j in range(4)]
        np.testing.assert_allclose(tj.bspline_basis(knots, 3, t), bern, atol=1e-14)


def test_spline_interpolates_log_poses():
    rng = np.random.default_rng(8)
    for kf in (4, 6):
        traj = tj.random_trajectory(rng, keyframes=kf)
        pts = np.stack([tj.se3_log(T) for T in traj.poses], axis=1)
        u = tj.chord_lengths(pts)
        for k in range(kf):
            np.testing.assert_allclose(traj(u[k]), traj.poses[k], atol=1e-10)


def test_chord_lengths_and_knots():
    pts = np.array([[0.0, 1.0, 3.0, 6.0]])
    np.testing.assert_allclose(tj.chord_lengths(pts), [0, 1 / 6, 3 / 6, 1])
    knots = tj.knot_averaging(np.array([0, 0.1, 0.2, 0.6, 1.0]), 3)
    np.testing.assert_allclose(knots, [0, 0, 0, 0, (0.1 + 0.2 + 0.6) / 3, 1, 1, 1, 1])


def test_deformation_util_conjugates_by_origin():
    rng = np.random.default_rng(2)
    xyz = rng.standard_normal((20, 3))
    handles = [3, 7, 11]
    O = tj.se3_exp(rng.standard_normal(6))
    T = tj.se3_exp(rng.standard_normal(6))
    du = tj.DeformationUtil(xyz, handles, origin=O)
    tabs = O @ T @ np.linalg.inv(O)
    exp = (tabs[:3, :3] @ xyz[handles].T).T + tabs[:3, 3]
    np.testing.assert_allclose(du.targets(T), exp, atol=1e-12)
    # identity motion keeps the handles in place
    np.testing.assert_allclose(du.targets(np.eye(4)), xyz[handles], atol=1e-12)


def test_cpp_header_matches_python_restatement():
    """include/deform/trajectory.h and trajectory.py are independent restatements of
    trajectory.h:61-73; on the reference's four poses they agree, and the C++ float path
    passes the reference's endpoint test."""
    import os
    import subprocess

    cpp = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
    subprocess.check_call(["make", "-s", "-C", cpp, "test_trajectory"])
    out = subprocess.run([os.path.join(cpp, "test_trajectory")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0] == "endpoints 1 1"
    path = tj.TrajectorySE3()
    pose0 = np.eye(4)
    path.add_key_pose(pose0)
    pose1 = translation(0, 0, 1) @ pose0
    path.add_key_pose(pose1)
    pose2 = translation(0, 0, 1) @ rotation_x(math.pi / 4) @ pose1
    path.add_key_pose(pose2)
    path.add_key_pose(translation(0, 0, 1) @ pose2)
    for line in lines[1:]:
        vals = [float(v) for v in line.split()[1:]]
        t, T = vals[0], np.array(vals[1:]).reshape(4, 4)
        np.testing.assert_allclose(T, path(t), atol=1e-12)


def test_cpp_transform_types_like_eigen():
    """trajectory.h and deformation_util.h with a transform type read as t.matrix()(r, c) (as
    Eigen::Transform is): the calls of tests/test_trajectory.cpp:17-38 and of
    examples/deform_trajectory.cpp:53-56,84 compile as written there, the endpoint test
    passes, and DeformationUtil's handle targets equal O T O^-1 p0."""
    import os
    import subprocess

    cpp = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp")
    subprocess.check_call(["make", "-s", "-C", cpp, "test_transform_types"])
    out = subprocess.run([os.path.join(cpp, "test_transform_types")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("transform types: ok")
