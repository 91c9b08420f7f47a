"""ctypes binding of libmdarap.so (include/md_arap.h) for the tests and the bench.

The library is the product: every call here goes through the C ABI to the HIP kernels.
There is no fallback: if the shared library or a GPU is missing, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdarap.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "md_arap.h")

MD_OK = 0
MD_ERR_INVALID, MD_ERR_NO_DEVICE, MD_ERR_HIP, MD_ERR_FACTOR = 1, 2, 3, 4
MD_ERR_NOMEM, MD_ERR_BUDGET, MD_ERR_STATE, MD_ERR_TIMEOUT = 5, 6, 7, 8
PREC_FP64, PREC_FP32_REFINE = 0, 1
SOLVER_LDLT, SOLVER_PCG = 0, 1
REST_REFERENCE, REST_FIXED = 0, 1

# candidate bottom-subtree budgets of the sweep-layout autotuner (LDS vector entries; 0: none)
AUTOTUNE_BUDGETS = (0, 128, 256, 512, 1024)
# candidate relaxed-amalgamation ratios of the factor-layout autotuner
AUTOTUNE_RATIOS = (0.3, 0.5, 0.7, 0.8, 0.9)

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


class Config(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("solver", ctypes.c_int), ("refine_steps", ctypes.c_int),
                ("rest_mode", ctypes.c_int), ("pcg_max_iters", ctypes.c_int), ("pcg_tol", ctypes.c_double),
                ("use_graph", ctypes.c_int), ("subtree_budget", ctypes.c_int), ("amalg_ratio", ctypes.c_double),
                ("factor_device", ctypes.c_int), ("level_pairs", ctypes.c_int)]


class MeshStats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nf", ctypes.c_int), ("F", ctypes.c_int), ("nslices", ctypes.c_int),
                ("supernodes", ctypes.c_int), ("levels", ctypes.c_int), ("max_sn_width", ctypes.c_int),
                ("max_sn_rows", ctypes.c_int), ("slots", ctypes.c_int64), ("cotan_nnz", ctypes.c_int64),
                ("panel_entries", ctypes.c_int64), ("factor_nnz", ctypes.c_int64),
                ("fwd_items", ctypes.c_int64), ("bwd_items", ctypes.c_int64),
                ("device_bytes", ctypes.c_int64), ("setup_ms_host", ctypes.c_double),
                ("bytes_local", ctypes.c_int64), ("bytes_residual", ctypes.c_int64),
                ("bytes_forward", ctypes.c_int64), ("bytes_backward", ctypes.c_int64),
                ("subtree_budget", ctypes.c_int), ("subtrees", ctypes.c_int), ("factor_nnz_exact", ctypes.c_int64),
                ("factor_ms", ctypes.c_double), ("deform_ms", ctypes.c_double * 6)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["deform_ms"] = list(self.deform_ms)
        return d


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_active", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("mesh_iterations", ctypes.c_uint64),
                ("evictions", ctypes.c_uint64), ("restores", ctypes.c_uint64),
                ("resident_bytes", ctypes.c_size_t), ("peak_resident_bytes", ctypes.c_size_t),
                ("meshes", ctypes.c_int), ("resident_meshes", ctypes.c_int), ("queued", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("md error %d: %s" % (code, msg))
        self.code = code


_lib = None

# (name, restype, argtypes) for every exported function
_SIGS = [
    ("md_config_default", None, [ctypes.POINTER(Config)]),
    ("md_last_error", ctypes.c_char_p, []),
    ("md_version", ctypes.c_char_p, []),
    ("md_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("md_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    ("md_ctx_destroy", ctypes.c_int, [_vp]),
    ("md_ctx_synchronize", ctypes.c_int, [_vp]),
    ("md_mesh_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p, ctypes.POINTER(Config),
                                      ctypes.POINTER(_vp)]),
    ("md_mesh_destroy", ctypes.c_int, [_vp]),
    ("md_mesh_set_constraint", ctypes.c_int, [_vp, ctypes.c_int, _f64p]),
    ("md_mesh_set_constraints", ctypes.c_int, [_vp, ctypes.c_int, _i32p, _f64p]),
    ("md_mesh_deform", ctypes.c_int, [_vp, ctypes.c_int, _f64p, ctypes.POINTER(ctypes.c_int)]),
    ("md_mesh_prepare", ctypes.c_int, [_vp, _f64p]),
    ("md_meshes_prepare", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int]),
    ("md_mesh_iterate", ctypes.c_int, [_vp, ctypes.c_int]),
    ("md_mesh_download", ctypes.c_int, [_vp, _f64p]),
    ("md_mesh_cotan_nnz", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("md_mesh_cotan_csr", ctypes.c_int, [_vp, _i32p, _i32p, _f64p]),
    ("md_mesh_rotations", ctypes.c_int, [_vp, _f64p]),
    ("md_mesh_normals", ctypes.c_int, [_vp, _f64p]),
    ("md_mesh_normals_noalloc", ctypes.c_int, [_vp, _f64p]),
    ("md_mesh_energy", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double)]),
    ("md_mesh_get_stats", ctypes.c_int, [_vp, ctypes.POINTER(MeshStats)]),
    ("md_mesh_time_iterations", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                               _f64p]),
    ("md_batch_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("md_batch_iterate", ctypes.c_int, [_vp, ctypes.c_int]),
    ("md_batch_time_iterations", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    ("md_mesh_autotune", ctypes.c_int, [_vp, _i32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_double)]),
    ("md_mesh_autotune_factor", ctypes.c_int, [_vp, _f64p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_double)]),
    ("md_batch_autotune", ctypes.c_int, [_vp, _i32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_double)]),
    ("md_batch_destroy", ctypes.c_int, [_vp]),
    ("md_server_create", ctypes.c_int, [_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    ("md_server_add_mesh", ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int)]),
    ("md_server_submit", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i32p, _f64p, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    ("md_server_poll", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), _vp,
                                      ctypes.POINTER(ctypes.c_double)]),
    ("md_server_wait", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_double, _vp, ctypes.POINTER(ctypes.c_double)]),
    ("md_server_get_stats", ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    ("md_server_destroy", ctypes.c_int, [_vp]),
    ("md_copy_probe", ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    ("md_host_best_rotation", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _f64p, _vp, _f64p, _vp]),
    ("md_host_admission_plan", ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _vp, _vp,
                                              ctypes.POINTER(ctypes.c_int), _vp, ctypes.POINTER(ctypes.c_int)]),
    ("md_host_factor_solve", ctypes.c_int, [ctypes.c_int, _i32p, _i32p, _f64p, _f64p, ctypes.c_int, _f64p, _f64p,
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]),
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libmdarap.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != MD_OK:
        raise MDError(rc, lib().md_last_error().decode())
    return rc


def exported_names():
    return [s[0] for s in _SIGS]


def device_count():
    c = ctypes.c_int(0)
    check(lib().md_device_count(ctypes.byref(c)))
    return c.value


def config(precision=PREC_FP64, rest_mode=REST_REFERENCE, refine_steps=1, use_graph=True, solver=SOLVER_LDLT,
           pcg_iters=None, pcg_tol=None, subtree_budget=None, amalg_ratio=None, factor_device=None,
           level_pairs=None):
    c = Config()
    lib().md_config_default(ctypes.byref(c))
    c.precision = precision
    c.rest_mode = rest_mode
    c.refine_steps = refine_steps
    c.use_graph = 1 if use_graph else 0
    c.solver = solver
    if pcg_iters is not None:
        c.pcg_max_iters = int(pcg_iters)
    if pcg_tol is not None:
        c.pcg_tol = float(pcg_tol)
    if subtree_budget is not None:
        c.subtree_budget = int(subtree_budget)
    if amalg_ratio is not None:
        c.amalg_ratio = float(amalg_ratio)
    if factor_device is not None:
        c.factor_device = 1 if factor_device else 0
    if level_pairs is not None:
        c.level_pairs = 1 if level_pairs else 0
    return c


class Context:
    def __init__(self, device=0):
        h = _vp()
        check(lib().md_ctx_create(int(device), ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        check(lib().md_ctx_synchronize(self._h))

    def close(self):
        if self._h:
            lib().md_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """The ARAP object (include/deform/arap.h's class through the C ABI)."""

    def __init__(self, ctx, xyz, tri, cfg=None):
        self.ctx = ctx
        self.tri = np.ascontiguousarray(tri, np.int32).reshape(-1, 3)
        self.mesh = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3).copy()
        self.n = self.mesh.shape[0]
        h = _vp()
        cfg = cfg if cfg is not None else config()
        check(lib().md_mesh_create(ctx.handle, self.n, self.tri.shape[0], self.tri.ravel(), ctypes.byref(cfg),
                                   ctypes.byref(h)))
        self._h = h
        self.cfg = cfg

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().md_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_constraint(self, vidx, xyz):
        check(lib().md_mesh_set_constraint(self._h, int(vidx), np.ascontiguousarray(xyz, np.float64).reshape(3)))

    def set_constraints(self, idx, xyz):
        idx = np.ascontiguousarray(idx, np.int32).ravel()
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        check(lib().md_mesh_set_constraints(self._h, idx.shape[0], idx, xyz.ravel()))

    def deform(self, iters):
        buf = self.mesh.ravel().copy()
        ok = ctypes.c_int(0)
        rc = lib().md_mesh_deform(self._h, int(iters), buf, ctypes.byref(ok))
        if rc == MD_ERR_FACTOR:
            return False
        check(rc)
        self.mesh = buf.reshape(-1, 3)
        return bool(ok.value)

    def prepare(self, xyz=None):
        buf = np.ascontiguousarray(self.mesh if xyz is None else xyz, np.float64).ravel().copy()
        check(lib().md_mesh_prepare(self._h, buf))

    def iterate(self, iters):
        check(lib().md_mesh_iterate(self._h, int(iters)))

    def download(self):
        out = np.zeros(3 * self.n, np.float64)
        check(lib().md_mesh_download(self._h, out))
        return out.reshape(-1, 3)

    def cotan(self):
        nnz = ctypes.c_int64(0)
        check(lib().md_mesh_cotan_nnz(self._h, ctypes.byref(nnz)))
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(max(nnz.value, 1), np.int32)
        v = np.zeros(max(nnz.value, 1), np.float64)
        check(lib().md_mesh_cotan_csr(self._h, rp, ci, v))
        return rp, ci[: nnz.value], v[: nnz.value]

    def rotations(self):
        R = np.zeros(9 * self.n, np.float64)
        check(lib().md_mesh_rotations(self._h, R))
        return R.reshape(self.n, 3, 3)

    def normals(self):
        out = np.zeros(3 * self.n, np.float64)
        check(lib().md_mesh_normals(self._h, out))
        return out.reshape(-1, 3)

    def normals_noalloc(self):
        """md_mesh_normals_noalloc: the same normals from the per-GPU static workspace."""
        out = np.zeros(3 * self.n, np.float64)
        check(lib().md_mesh_normals_noalloc(self._h, out))
        return out.reshape(-1, 3)

    def energy(self):
        e = ctypes.c_double(0)
        check(lib().md_mesh_energy(self._h, ctypes.byref(e)))
        return e.value

    def stats(self):
        s = MeshStats()
        check(lib().md_mesh_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def time_iterations(self, warmup, iters):
        tot = ctypes.c_double(0)
        ph = np.zeros(4, np.float64)
        check(lib().md_mesh_time_iterations(self._h, int(warmup), int(iters), ctypes.byref(tot), ph))
        return tot.value, ph

    def autotune(self, budgets=AUTOTUNE_BUDGETS, iters=5):
        """md_mesh_autotune: keeps the fastest sweep layout; returns (budget, ms per iteration)."""
        return _autotune(lib().md_mesh_autotune, self._h, budgets, iters)

    def autotune_factor(self, ratios=AUTOTUNE_RATIOS, budgets=AUTOTUNE_BUDGETS, iters=5):
        """md_mesh_autotune_factor (freshly prepared mesh): fastest (amalgamation ratio, subtree
        budget); returns (ratio, budget, ms per iteration)."""
        r = np.ascontiguousarray(ratios, np.float64)
        b = np.ascontiguousarray(budgets, np.int32)
        br, bb, ms = ctypes.c_double(0), ctypes.c_int(0), ctypes.c_double(0)
        check(lib().md_mesh_autotune_factor(self._h, r, len(r), b, len(b), int(iters), ctypes.byref(br),
                                            ctypes.byref(bb), ctypes.byref(ms)))
        return br.value, bb.value, ms.value


def _autotune(fn, h, budgets, iters):
    b = np.ascontiguousarray(budgets, np.int32)
    best = ctypes.c_int(0)
    ms = ctypes.c_double(0)
    check(fn(h, b, len(b), int(iters), ctypes.byref(best), ctypes.byref(ms)))
    return best.value, ms.value


def prepare_many(ctx, meshes, positions=None, nthreads=0):
    """md_meshes_prepare: the dirty setup of many meshes, host work on parallel threads."""
    bufs = [np.ascontiguousarray(m.mesh if positions is None else positions[k], np.float64).ravel().copy()
            for k, m in enumerate(meshes)]
    hs = (_vp * len(meshes))(*[m.handle for m in meshes])
    ps = (_vp * len(meshes))(*[b.ctypes.data for b in bufs])
    check(lib().md_meshes_prepare(ctx.handle, len(meshes), hs, ps, int(nthreads)))


class Batch:
    def __init__(self, ctx, meshes):
        self.ctx = ctx
        self.meshes = list(meshes)
        arr = (_vp * len(self.meshes))(*[m.handle for m in self.meshes])
        h = _vp()
        check(lib().md_batch_create(ctx.handle, len(self.meshes), arr, ctypes.byref(h)))
        self._h = h

    def iterate(self, iters):
        check(lib().md_batch_iterate(self._h, int(iters)))

    def time_iterations(self, warmup, iters):
        tot = ctypes.c_double(0)
        check(lib().md_batch_time_iterations(self._h, int(warmup), int(iters), ctypes.byref(tot)))
        return tot.value

    def autotune(self, budgets=AUTOTUNE_BUDGETS, iters=3):
        """md_batch_autotune: one sweep layout for every mesh; returns (budget, ms per iteration)."""
        return _autotune(lib().md_batch_autotune, self._h, budgets, iters)

    def close(self):
        if getattr(self, "_h", None):
            lib().md_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    def __init__(self, ctx, hbm_budget_bytes=0, max_active=0):
        self.ctx = ctx
        c = ServerConfig()
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.max_active = int(max_active)
        h = _vp()
        check(lib().md_server_create(ctx.handle, ctypes.byref(c), ctypes.byref(h)))
        self._h = h
        self.meshes = {}

    def add_mesh(self, mesh):
        mid = ctypes.c_int(-1)
        check(lib().md_server_add_mesh(self._h, mesh.handle, ctypes.byref(mid)))
        self.meshes[mid.value] = mesh
        return mid.value

    def submit(self, mesh_id, idx, xyz, iters):
        idx = np.ascontiguousarray(idx, np.int32).ravel()
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        t = ctypes.c_uint64(0)
        check(lib().md_server_submit(self._h, int(mesh_id), idx.shape[0], idx, xyz.ravel(), int(iters),
                                     ctypes.byref(t)))
        return t.value

    def poll(self, ticket, n=None):
        done = ctypes.c_int(0)
        lat = ctypes.c_double(0)
        out = np.zeros(3 * n, np.float64) if n else None
        check(lib().md_server_poll(self._h, ctypes.c_uint64(ticket), ctypes.byref(done),
                                   out.ctypes.data if out is not None else None, ctypes.byref(lat)))
        if not done.value:
            return None
        return (out.reshape(-1, 3) if out is not None else None), lat.value

    def wait(self, ticket, n=None, timeout_ms=60000.0):
        lat = ctypes.c_double(0)
        out = np.zeros(3 * n, np.float64) if n else None
        check(lib().md_server_wait(self._h, ctypes.c_uint64(ticket), float(timeout_ms),
                                   out.ctypes.data if out is not None else None, ctypes.byref(lat)))
        return (out.reshape(-1, 3) if out is not None else None), lat.value

    def stats(self):
        st = ServerStats()
        check(lib().md_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().md_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_probe(ctx, nbytes, repeats=3):
    """md_copy_probe: plain device copy of nbytes, repeats times; achieved GB/s (read + write)."""
    g = ctypes.c_double(0)
    check(lib().md_copy_probe(ctx.handle, int(nbytes), int(repeats), ctypes.byref(g)))
    return g.value


def host_factor_solve(rp, ci, v, xyz, b, subtree_budget=-1, device_factor_emulation=False, level_pairs=False):
    """Host-only: the library's ordering + symbolic + FP64 block LDL^T, solved on the CPU by an
    emulation of the device sweeps (subtree_budget as in Config; -1: the default); with
    device_factor_emulation the factorization is the emulation of the device kernels."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    b2 = np.atleast_2d(np.asarray(b, np.float64))
    n = rp.shape[0] - 1
    x = np.zeros_like(b2)
    pe = ctypes.c_int64(0)
    lv = ctypes.c_int(0)
    check(lib().md_host_factor_solve(n, rp, ci, v, np.ascontiguousarray(xyz, np.float64).ravel(), b2.shape[0],
                                     np.ascontiguousarray(b2).ravel(), x.reshape(-1), ctypes.byref(pe),
                                     ctypes.byref(lv), int(subtree_budget), 1 if device_factor_emulation else 0,
                                     1 if level_pairs else 0))
    return (x[0] if np.ndim(b) == 1 else x), pe.value, lv.value


def host_best_rotation(cov, prev=None, fp64=True):
    """Host-only: the local step's rotation code on the CPU for covariances (k, 3, 3); returns
    (quaternions (k, 4) as x, y, z, w, per-entry flag of the Jacobi fallback)."""
    cov = np.ascontiguousarray(cov, np.float64).reshape(-1, 9)
    k = cov.shape[0]
    q = np.zeros((k, 4), np.float64)
    fb = np.zeros(k, np.int32)
    pv = None if prev is None else np.ascontiguousarray(prev, np.float64).reshape(k, 4)
    check(lib().md_host_best_rotation(1 if fp64 else 0, k, cov.ravel(), None if pv is None else pv.ctypes.data,
                                      q.reshape(-1), fb.ctypes.data))
    return q, fb.astype(bool)


def host_admission_plan(bytes_, last_use, resident, busy, queue, active=0, max_active=0, budget=0,
                        resident_bytes=None):
    """Host-only: one admission pass of the server's planner. Returns (take per queued request,
    meshes to evict, meshes to restore, resident bytes after both)."""
    b = np.ascontiguousarray(bytes_, np.int64)
    lu = np.ascontiguousarray(last_use, np.uint64)
    rs = np.ascontiguousarray(resident, np.int32)
    bz = np.ascontiguousarray(busy, np.int32)
    q = np.ascontiguousarray(queue, np.int32)
    n = b.shape[0]
    take = np.zeros(max(1, q.shape[0]), np.int32)
    ev = np.zeros(max(1, n), np.int32)
    re = np.zeros(max(1, n), np.int32)
    nev, nre = ctypes.c_int(0), ctypes.c_int(0)
    rb = ctypes.c_int64(int(b[rs != 0].sum()) if resident_bytes is None else int(resident_bytes))
    check(lib().md_host_admission_plan(n, b.ctypes.data, lu.ctypes.data, rs.ctypes.data, bz.ctypes.data, q.shape[0],
                                       q.ctypes.data, int(active), int(max_active), int(budget), ctypes.byref(rb),
                                       take.ctypes.data, ev.ctypes.data, ctypes.byref(nev), re.ctypes.data,
                                       ctypes.byref(nre)))
    return take[: q.shape[0]].astype(bool), ev[: nev.value].tolist(), re[: nre.value].tolist(), rb.value


def quat_to_matrix(q):
    """Rotation matrices (k, 3, 3) of quaternions (k, 4) as x, y, z, w."""
    q = np.asarray(q, np.float64).reshape(-1, 4)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    s = 2.0 / (x * x + y * y + z * z + w * w)
    return np.stack([1 - s * (y * y + z * z), s * (x * y - w * z), s * (x * z + w * y),
                     s * (x * y + w * z), 1 - s * (x * x + z * z), s * (y * z - w * x),
                     s * (x * z - w * y), s * (y * z + w * x), 1 - s * (x * x + y * y)], axis=1).reshape(-1, 3, 3)
