"""Deterministic synthetic meshes and constraint sets for the ARAP workloads.

Every generator is pure numpy and seeded with ``numpy.random.default_rng(seed)``, so the
same call gives the same arrays on every machine. Meshes are returned as
``(xyz float64 (n,3), tri int32 (F,3))`` with counter-clockwise (outward) faces, the
convention the reference's adapter asks for (inc/deform/openmesh_adapter.h:86-99).

Workloads (vertex and face counts are fixed by construction):
  * ``uv_sphere(40, 80)``: r=1, 40 interior parallels at theta_k = k*pi/41 plus 2 poles,
    n=3,202, F=6,400; anchors z<-0.8, handles z>0.8 moved by (0.3,0,0.5).
  * ``lattice_bar(400, 40, 40)``: surface of the box [0,2]x[-0.1,0.1]^2 on a 400x40x40
    lattice, each quad split along the same diagonal, n=67,202, F=134,400; anchors are the
    lattice columns i<=9 (x<0.05), handles i>=391 (x>1.95), handles rotated 90 degrees about
    the x axis and moved by (0,0.3,0).
  * ``icosphere(7)`` with seeded radial noise sigma=0.01: n=163,842, F=327,680; anchors are
    the lowest floor(n/10) vertices by z, handles the highest floor(n/20), lifted by 0.4 in z.
  * ``mixed_batch``: meshes with log-uniform target sizes (icospheres of the nearest level or
    lattice bars sized to the target), per-mesh anchors 5-15% and handles 1-5% by z
    quantiles, handle offsets uniform in [-0.5,0.5]^3.
"""
import math

import numpy as np


# ------------------------------------------------------------------------------------------
# meshes
# ------------------------------------------------------------------------------------------

def uv_sphere(rings=40, segs=80, r=1.0):
    """UV sphere: north pole, ``rings`` parallels at theta_k = k*pi/(rings+1), south pole."""
    th = np.arange(1, rings + 1) * math.pi / (rings + 1)
    ph = np.arange(segs) * 2.0 * math.pi / segs
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    xyz = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]]) * r
    n = xyz.shape[0]
    south = n - 1

    def vid(k, j):  # ring k in [0, rings), segment j
        return 1 + k * segs + (j % segs)

    faces = []
    j = np.arange(segs)
    faces.append(np.stack([np.zeros(segs, np.int64), vid(0, j), vid(0, j + 1)], -1))
    for k in range(rings - 1):
        a, b = vid(k, j), vid(k, j + 1)
        c, d = vid(k + 1, j), vid(k + 1, j + 1)
        faces.append(np.stack([a, c, d], -1))
        faces.append(np.stack([a, d, b], -1))
    faces.append(np.stack([np.full(segs, south), vid(rings - 1, j + 1), vid(rings - 1, j)], -1))
    tri = np.concatenate(faces).astype(np.int32)
    return xyz.astype(np.float64), tri


def lattice_bar(nx=400, ny=40, nz=40, lo=(0.0, -0.1, -0.1), hi=(2.0, 0.1, 0.1)):
    """Surface of an axis-aligned box on an (nx, ny, nz) cell lattice.

    Returns (xyz, tri, lattice_i) where lattice_i is the x lattice index of each vertex.
    Each surface quad is split along its (0,0)-(1,1) diagonal.
    """
    gx, gy, gz = nx + 1, ny + 1, nz + 1
    I, J, K = np.meshgrid(np.arange(gx), np.arange(gy), np.arange(gz), indexing="ij")
    on = (I == 0) | (I == nx) | (J == 0) | (J == ny) | (K == 0) | (K == nz)
    idx = -np.ones((gx, gy, gz), np.int64)
    idx[on] = np.arange(int(on.sum()))
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    step = (hi - lo) / np.array([nx, ny, nz], np.float64)
    xyz = np.stack([lo[0] + I[on] * step[0], lo[1] + J[on] * step[1], lo[2] + K[on] * step[2]], -1)
    lat_i = I[on].astype(np.int32)

    faces = []

    def quads(a, b, c, d, flip):
        # quad corners a=(0,0) b=(1,0) c=(1,1) d=(0,1); outward orientation chosen by flip
        a, b, c, d = (x.ravel() for x in (a, b, c, d))
        if flip:
            faces.append(np.stack([a, c, b], -1))
            faces.append(np.stack([a, d, c], -1))
        else:
            faces.append(np.stack([a, b, c], -1))
            faces.append(np.stack([a, c, d], -1))

    # x = 0 and x = nx faces (grid over j, k)
    for i, flip in ((0, True), (nx, False)):
        s = idx[i]
        quads(s[:-1, :-1], s[1:, :-1], s[1:, 1:], s[:-1, 1:], flip)
    # y = 0 and y = ny faces (grid over k, i)
    for jj, flip in ((0, True), (ny, False)):
        s = idx[:, jj, :].T  # (k, i)
        quads(s[:-1, :-1], s[1:, :-1], s[1:, 1:], s[:-1, 1:], flip)
    # z = 0 and z = nz faces (grid over i, j)
    for kk, flip in ((0, True), (nz, False)):
        s = idx[:, :, kk]  # (i, j)
        quads(s[:-1, :-1], s[1:, :-1], s[1:, 1:], s[:-1, 1:], flip)
    tri = np.concatenate(faces).astype(np.int32)
    return xyz, tri, lat_i


_ICO_CACHE = {}


def icosphere(level):
    """Unit icosphere by ``level`` midpoint subdivisions: n = 10*4^level + 2."""
    if level in _ICO_CACHE:
        x, t = _ICO_CACHE[level]
        return x.copy(), t.copy()
    g = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0],
                  [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        mids = (v[uk // nv] + v[uk % nv]) * 0.5
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv  # m[0]: edge (0,1), m[1]: (1,2), m[2]: (2,0)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], -1), np.stack([b, bc, ab], -1),
                            np.stack([c, ca, bc], -1), np.stack([ab, bc, ca], -1)])
        v = np.concatenate([v, mids])
    tri = f.astype(np.int32)
    _ICO_CACHE[level] = (v, tri)
    return v.copy(), tri.copy()


def radial_noise(xyz, sigma, seed):
    rng = np.random.default_rng(seed)
    s = 1.0 + sigma * rng.standard_normal(xyz.shape[0])
    return xyz * s[:, None]


def z_order(xyz):
    """Vertex indices sorted by z, ties by index (stable)."""
    return np.lexsort((np.arange(xyz.shape[0]), xyz[:, 2]))


def rot_x(angle):
    c, s = math.cos(angle), math.sin(angle)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float64)


# ------------------------------------------------------------------------------------------
# workloads
# ------------------------------------------------------------------------------------------

class Problem:
    """A mesh plus its constraint set: anchors keep their rest position, handles move."""

    def __init__(self, name, xyz, tri, anchors, handles, handle_targets):
        self.name = name
        self.xyz = np.ascontiguousarray(xyz, np.float64)
        self.tri = np.ascontiguousarray(tri, np.int32)
        self.anchors = np.asarray(anchors, np.int32)
        self.handles = np.asarray(handles, np.int32)
        self.handle_targets = np.ascontiguousarray(handle_targets, np.float64).reshape(-1, 3)

    @property
    def n(self):
        return self.xyz.shape[0]

    @property
    def F(self):
        return self.tri.shape[0]

    def constraints(self):
        """(indices, targets) with anchors first, then handles."""
        idx = np.concatenate([self.anchors, self.handles]).astype(np.int32)
        tgt = np.concatenate([self.xyz[self.anchors], self.handle_targets])
        return idx, np.ascontiguousarray(tgt)


def sphere_problem(rings=40, segs=80):
    xyz, tri = uv_sphere(rings, segs)
    anchors = np.nonzero(xyz[:, 2] < -0.8)[0]
    handles = np.nonzero(xyz[:, 2] > 0.8)[0]
    tgt = xyz[handles] + np.array([0.3, 0.0, 0.5])
    return Problem("uv_sphere_%dx%d" % (rings, segs), xyz, tri, anchors, handles, tgt)


def bar_problem(nx=400, ny=40, nz=40):
    xyz, tri, li = lattice_bar(nx, ny, nz)
    edge = max(1, int(round(nx * 0.05 / 2.0)))  # lattice columns with x < 0.05 (exclusive)
    anchors = np.nonzero(li < edge)[0]
    handles = np.nonzero(li > nx - edge)[0]
    tgt = xyz[handles] @ rot_x(math.pi / 2).T + np.array([0.0, 0.3, 0.0])
    return Problem("lattice_bar_%dx%dx%d" % (nx, ny, nz), xyz, tri, anchors, handles, tgt)


def noisy_icosphere_problem(level=7, sigma=0.01, seed=1234):
    xyz, tri = icosphere(level)
    xyz = radial_noise(xyz, sigma, seed)
    n = xyz.shape[0]
    order = z_order(xyz)
    anchors = np.sort(order[: n // 10])
    handles = np.sort(order[n - n // 20:])
    tgt = xyz[handles] + np.array([0.0, 0.0, 0.4])
    return Problem("icosphere%d_noise%g_seed%d" % (level, sigma, seed), xyz, tri, anchors, handles, tgt)


def _bar_for_target(n_target):
    # lattice bar with cells (10b, b, b): n = 42 b^2 + 2 approximately
    b = max(2, int(round(math.sqrt(n_target / 42.0))))
    xyz, tri, _ = lattice_bar(10 * b, b, b, lo=(0.0, -0.1, -0.1), hi=(2.0, 0.1, 0.1))
    return xyz, tri


def mixed_problem(k, n_target, rng):
    if k % 2 == 0:
        level = int(np.clip(round(math.log(max(n_target - 2, 12) / 10.0, 4)), 2, 7))
        xyz, tri = icosphere(level)
        xyz = radial_noise(xyz, 0.01, int(rng.integers(1 << 30)))
        kind = "icosphere%d" % level
    else:
        xyz, tri = _bar_for_target(n_target)
        kind = "bar"
    n = xyz.shape[0]
    fa = rng.uniform(0.05, 0.15)
    fh = rng.uniform(0.01, 0.05)
    order = z_order(xyz)
    na, nh = max(1, int(fa * n)), max(1, int(fh * n))
    anchors = np.sort(order[:na])
    handles = np.sort(order[n - nh:])
    off = rng.uniform(-0.5, 0.5, size=3)
    tgt = xyz[handles] + off
    return Problem("mix%03d_%s_n%d" % (k, kind, n), xyz, tri, anchors, handles, tgt)


def mixed_batch(count=256, lo=1000, hi=200000, seed=4242):
    """``count`` meshes with log-uniform target sizes in [lo, hi]."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))
    return [mixed_problem(k, int(t), rng) for k, t in enumerate(targets)]


def relative_error(a, b):
    """max |a - b| / max |b| over all coordinates (the position-error measure)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
