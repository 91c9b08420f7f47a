"""ctypes view of the FP64 CPU oracle (oracle/arap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``. The product path never imports it.

The oracle restates inc/deform/arap.h of cheind/mesh-deform (file:line citations are in the
C source next to each function).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libarap_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_cotan_csr.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, _i32p, _i32p, _f64p, ctypes.c_int]
        L.or_cotan_csr.restype = ctypes.c_int
        L.or_svd_rotation.argtypes = [_f64p, _f64p]
        L.or_svd3.argtypes = [_f64p, _f64p, _f64p, _f64p]
        L.or_nd_order.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _i32p]
        L.or_nd_order.restype = ctypes.c_int
        L.or_spd_solve.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_void_p, ctypes.c_int,
                                   _f64p, _f64p, ctypes.POINTER(ctypes.c_longlong)]
        L.or_spd_solve.restype = ctypes.c_int
        L.or_arap_create.argtypes = [ctypes.c_int, ctypes.c_int, _i32p]
        L.or_arap_create.restype = ctypes.c_void_p
        L.or_arap_destroy.argtypes = [ctypes.c_void_p]
        L.or_arap_set_constraint.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p]
        L.or_arap_set_constraint.restype = ctypes.c_int
        L.or_arap_deform.argtypes = [ctypes.c_void_p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_void_p]
        L.or_arap_deform.restype = ctypes.c_int
        L.or_arap_energy.argtypes = [ctypes.c_void_p]
        L.or_arap_energy.restype = ctypes.c_double
        L.or_arap_nnz_weights.argtypes = [ctypes.c_void_p]
        L.or_arap_nnz_weights.restype = ctypes.c_int
        L.or_arap_cotan.argtypes = [ctypes.c_void_p, _i32p, _i32p, _f64p]
        L.or_arap_cotan.restype = ctypes.c_int
        for name in ("or_arap_rotations", "or_arap_pprime", "or_arap_rest", "or_arap_bfixed", "or_arap_rhs"):
            getattr(L, name).argtypes = [ctypes.c_void_p, _f64p]
        L.or_arap_nfree.argtypes = [ctypes.c_void_p]
        L.or_arap_nfree.restype = ctypes.c_int
        L.or_arap_factor_nnz.argtypes = [ctypes.c_void_p]
        L.or_arap_factor_nnz.restype = ctypes.c_longlong
        L.or_arap_laplacian_nnz.argtypes = [ctypes.c_void_p]
        L.or_arap_laplacian_nnz.restype = ctypes.c_int
        L.or_rotations.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def cotan_csr(xyz, tri):
    """Cotan weight matrix of arap.h:182-239 as (rowptr, col, val)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 3)
    n, F = xyz.shape[0], tri.shape[0]
    cap = 6 * F + 1
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap, np.float64)
    nnz = lib().or_cotan_csr(n, F, tri.ravel(), xyz.ravel(), rp, ci, v, cap)
    if nnz < 0:
        raise ValueError("or_cotan_csr failed: %d" % nnz)
    return rp, ci[:nnz].copy(), v[:nnz].copy()


def svd_rotation(cov):
    """R = V diag(1,1,det(V U^T)) U^T of arap.h:376-382 for a 3x3 covariance."""
    c = np.ascontiguousarray(cov, dtype=np.float64).reshape(9)
    R = np.zeros(9, np.float64)
    lib().or_svd_rotation(c, R)
    return R.reshape(3, 3)


def svd3(A):
    a = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    U, S, V = np.zeros(9), np.zeros(3), np.zeros(9)
    lib().or_svd3(a, U, S, V)
    return U.reshape(3, 3), S, V.reshape(3, 3)


def rotations(rp, ci, w, p, pp):
    n = p.shape[0]
    R = np.zeros(9 * n, np.float64)
    lib().or_rotations(n, np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
                       np.ascontiguousarray(w, np.float64), np.ascontiguousarray(p, np.float64).ravel(),
                       np.ascontiguousarray(pp, np.float64).ravel(), R)
    return R.reshape(n, 3, 3)


def spd_solve(rp, ci, v, b, xyz=None):
    """Solve A x = b with the oracle's sparse LDL^T; b is (n,) or (k, n)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b2 = np.atleast_2d(np.asarray(b, np.float64))
    n = rp.shape[0] - 1
    x = np.zeros_like(b2)
    nnzl = ctypes.c_longlong(0)
    xyz_arr = None if xyz is None else np.ascontiguousarray(xyz, np.float64).ravel()
    rc = lib().or_spd_solve(n, rp, ci, v, None if xyz_arr is None else xyz_arr.ctypes.data,
                            b2.shape[0], np.ascontiguousarray(b2).ravel(), x.reshape(-1), ctypes.byref(nnzl))
    if rc != 0:
        raise RuntimeError("or_spd_solve failed: %d" % rc)
    return (x[0] if np.ndim(b) == 1 else x), nnzl.value


class Arap:
    """The oracle ARAP object; mirrors AsRigidAsPossibleDeformation (arap.h:49-466).

    ``mesh`` is an (n, 3) float64 array standing in for the mesh adapter: it is read on
    dirty deforms and overwritten by the write-back, exactly like the adapter.
    """

    def __init__(self, mesh_xyz, tri):
        self.tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        self.mesh = np.ascontiguousarray(mesh_xyz, np.float64).reshape(-1, 3).copy()
        self.n = self.mesh.shape[0]
        self._h = lib().or_arap_create(self.n, self.tri.shape[0], self.tri.ravel())
        if not self._h:
            raise ValueError("or_arap_create failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_arap_destroy(h)
            self._h = None

    def set_constraint(self, vidx, xyz):
        rc = lib().or_arap_set_constraint(self._h, int(vidx), np.ascontiguousarray(xyz, np.float64).reshape(3))
        if rc != 0:
            raise IndexError("constraint index out of range: %d" % vidx)

    def set_constraints(self, idx, xyz):
        xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
        for i, v in zip(np.asarray(idx).ravel(), xyz):
            self.set_constraint(int(i), v)

    def deform(self, iters, rest_mode=0, energy=False):
        buf = np.ascontiguousarray(self.mesh).ravel().copy()
        en = np.zeros(max(iters, 1), np.float64)
        rc = lib().or_arap_deform(self._h, int(iters), buf, int(rest_mode), en.ctypes.data if energy else None)
        if rc < 0:
            raise RuntimeError("or_arap_deform failed: %d" % rc)
        self.mesh = buf.reshape(-1, 3)
        if energy:
            return bool(rc), en[:iters].copy()
        return bool(rc)

    def energy(self):
        return lib().or_arap_energy(self._h)

    def cotan(self):
        nnz = lib().or_arap_nnz_weights(self._h)
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1), np.float64)
        lib().or_arap_cotan(self._h, rp, ci, v)
        return rp, ci[:nnz], v[:nnz]

    def rotations(self):
        R = np.zeros(9 * self.n, np.float64)
        lib().or_arap_rotations(self._h, R)
        return R.reshape(self.n, 3, 3)

    def pprime(self):
        out = np.zeros(3 * self.n, np.float64)
        lib().or_arap_pprime(self._h, out)
        return out.reshape(self.n, 3)

    def rest(self):
        out = np.zeros(3 * self.n, np.float64)
        lib().or_arap_rest(self._h, out)
        return out.reshape(self.n, 3)

    def nfree(self):
        return lib().or_arap_nfree(self._h)

    def bfixed(self):
        nf = self.nfree()
        out = np.zeros(max(3 * nf, 1), np.float64)
        lib().or_arap_bfixed(self._h, out)
        return out[: 3 * nf].reshape(nf, 3)

    def rhs(self):
        nf = self.nfree()
        out = np.zeros(max(3 * nf, 1), np.float64)
        lib().or_arap_rhs(self._h, out)
        return out[: 3 * nf].reshape(nf, 3)

    def factor_nnz(self):
        return lib().or_arap_factor_nnz(self._h)

    def laplacian_nnz(self):
        return lib().or_arap_laplacian_nnz(self._h)
