"""Handle motion on the host: SE3 keyframe trajectories and handle-target generation.

Restates inc/deform/trajectory.h (TrajectorySE3, trajectory.h:33-83) and
inc/deform/deformation_util.h (DeformationUtil, deformation_util.h:19-64) of
cheind/mesh-deform. The reference builds on Sophus (SE3 log/exp) and Eigen's
SplineFitting; neither is available here, so both are restated from their published
algorithms:

* SE3 log/exp with tangent (upsilon, omega): omega is the rotation vector of R, t = V upsilon
  with V = I + (1 - cos th)/th^2 W + (th - sin th)/th^3 W^2, W = [omega]_x (Sophus SE3Group).
* Eigen::SplineFitting<Spline<S,6,3>>::Interpolate(points, 3): parameters by normalised
  chord lengths (SplineFitting::ChordLengths), knots by averaging (KnotAveraging: clamped,
  interior knot j+p = mean of parameters j .. j+p-1), control points from the collocation
  system sum_j N_j(u_i) P_j = points_i with Cox-de Boor B-spline bases. With 4 keyframes
  the knot vector is [0,0,0,0,1,1,1,1]: one cubic Bezier segment through the 4 log-poses.

Host-side only: these produce handle targets for the served drag stream; no kernel runs here.
"""
import math

import numpy as np


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def so3_exp(w):
    th = float(np.linalg.norm(w))
    W = hat(w)
    if th < 1e-10:
        return np.eye(3) + W + 0.5 * W @ W
    return np.eye(3) + math.sin(th) / th * W + (1 - math.cos(th)) / (th * th) * W @ W


def so3_log(R):
    c = (np.trace(R) - 1.0) * 0.5
    c = min(1.0, max(-1.0, c))
    th = math.acos(c)
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    if th < 1e-7:
        return 0.5 * v
    if math.pi - th < 1e-6:
        # axis from the symmetric part near pi
        A = (R + np.eye(3)) * 0.5
        k = int(np.argmax(np.diag(A)))
        axis = A[:, k] / math.sqrt(max(A[k, k], 1e-300))
        return axis * th
    return th / (2.0 * math.sin(th)) * v


def se3_exp(xi):
    """xi = (upsilon, omega) -> 4x4 matrix."""
    ups, w = np.asarray(xi[:3], float), np.asarray(xi[3:], float)
    th = float(np.linalg.norm(w))
    W = hat(w)
    if th < 1e-10:
        V = np.eye(3) + 0.5 * W + W @ W / 6.0
    else:
        V = np.eye(3) + (1 - math.cos(th)) / th ** 2 * W + (th - math.sin(th)) / th ** 3 * W @ W
    T = np.eye(4)
    T[:3, :3] = so3_exp(w)
    T[:3, 3] = V @ ups
    return T


def se3_log(T):
    R, t = T[:3, :3], T[:3, 3]
    w = so3_log(R)
    th = float(np.linalg.norm(w))
    W = hat(w)
    if th < 1e-10:
        Vinv = np.eye(3) - 0.5 * W + W @ W / 12.0
    else:
        Vinv = np.eye(3) - 0.5 * W + (1 - th / (2 * math.tan(th / 2))) / th ** 2 * W @ W
    return np.concatenate([Vinv @ t, w])


def chord_lengths(pts):
    """Normalised cumulative chord lengths of the columns of pts (d x n)."""
    n = pts.shape[1]
    u = np.zeros(n)
    if n > 1:
        u[1:] = np.cumsum(np.linalg.norm(np.diff(pts, axis=1), axis=0))
        u = u / u[-1] if u[-1] > 0 else np.linspace(0.0, 1.0, n)
        u[-1] = 1.0
    return u


def knot_averaging(u, p):
    n = u.shape[0]
    knots = np.zeros(n + p + 1)
    knots[n:] = 1.0
    for j in range(1, n - p):
        knots[j + p] = np.mean(u[j:j + p])
    return knots


def bspline_basis(knots, p, u):
    """Values of all n = len(knots)-p-1 basis functions of degree p at u (Cox-de Boor)."""
    n = len(knots) - p - 1
    # degree 0 (u = 1 belongs to the last non-empty span)
    N = np.zeros(len(knots) - 1)
    if u >= knots[-1]:
        span = np.max(np.nonzero(knots[:-1] < knots[-1])[0])
        N[span] = 1.0
    else:
        for i in range(len(knots) - 1):
            if knots[i] <= u < knots[i + 1]:
                N[i] = 1.0
    for d in range(1, p + 1):
        M = np.zeros(len(knots) - 1 - d)
        for i in range(len(M)):
            a = 0.0
            den = knots[i + d] - knots[i]
            if den > 0:
                a += (u - knots[i]) / den * N[i]
            den = knots[i + d + 1] - knots[i + 1]
            if den > 0:
                a += (knots[i + d + 1] - u) / den * N[i + 1]
            M[i] = a
        N = M
    return N[:n]


class TrajectorySE3:
    """Keyframe trajectory interpolated in se(3) (trajectory.h:33-83)."""

    def __init__(self):
        self.poses = []
        self._dirty = False
        self._ctrl = None
        self._knots = None
        self.degree = 3

    def add_key_pose(self, T):
        """addKeyPose (trajectory.h:55-59); returns the pose."""
        T = np.asarray(T, float).copy()
        self.poses.append(T)
        self._dirty = True
        return T

    def _fit(self):
        pts = np.stack([se3_log(T) for T in self.poses], axis=1)  # 6 x n (trajectory.h:63-66)
        p = min(self.degree, pts.shape[1] - 1)
        u = chord_lengths(pts)
        knots = knot_averaging(u, p)
        A = np.stack([bspline_basis(knots, p, ui) for ui in u])  # collocation matrix
        self._ctrl = np.linalg.solve(A, pts.T)  # n x 6 control points (trajectory.h:67)
        self._knots = knots
        self._p = p
        self._dirty = False

    def __call__(self, t):
        """operator()(time) (trajectory.h:61-73): exp(spline(t))."""
        if self._dirty:
            self._fit()
        N = bspline_basis(self._knots, self._p, float(t))
        return se3_exp(N @ self._ctrl)


class DeformationUtil:
    """Moves a captured handle set rigidly (deformation_util.h:19-64)."""

    def __init__(self, xyz, handles, origin=None):
        self.handles = np.asarray(handles, np.int64)
        self.points = np.asarray(xyz, float)[self.handles].copy()  # deformation_util.h:39-43
        self.origin = np.eye(4) if origin is None else np.asarray(origin, float)
        R, t = self.origin[:3, :3], self.origin[:3, 3]
        self.origin_inv = np.eye(4)  # isometry inverse (deformation_util.h:38)
        self.origin_inv[:3, :3] = R.T
        self.origin_inv[:3, 3] = -R.T @ t

    def targets(self, T):
        """The targets updateConstraints would set (deformation_util.h:48-57)."""
        tabs = self.origin @ np.asarray(T, float) @ self.origin_inv
        return self.points @ tabs[:3, :3].T + tabs[:3, 3]

    def update_constraints(self, T, arap):
        arap.set_constraints(self.handles, self.targets(T))


def random_trajectory(rng, keyframes=4, max_angle=0.6, max_shift=0.3):
    """Seeded 6-DoF keyframe trajectory starting at the identity."""
    traj = TrajectorySE3()
    T = np.eye(4)
    traj.add_key_pose(T)
    for _ in range(keyframes - 1):
        w = rng.uniform(-1, 1, size=3)
        w *= rng.uniform(0, max_angle) / max(np.linalg.norm(w), 1e-12)
        step = np.eye(4)
        step[:3, :3] = so3_exp(w)
        step[:3, 3] = rng.uniform(-max_shift, max_shift, size=3)
        T = T @ step
        traj.add_key_pose(T)
    return traj
