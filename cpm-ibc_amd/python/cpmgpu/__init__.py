"""ctypes binding of libcpmgpu.so (the MI355X closest-point engine's C ABI, include/cpm_gpu.h).

This is the Python side of the drop-in boundary used by the tests and the
benchmark.  It never falls back to a CPU path: if the HIP library is missing
or no GPU is visible, every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "..", "libcpmgpu.so"))

CPM_FP64 = 0
CPM_FP32_REFINE = 1
PRECOND_IDENTITY, PRECOND_DIAGONAL, PRECOND_JACOBI = 0, 1, 2
STATUS = {1: "converged", 2: "maxit", 4: "zero_rhs", 5: "breakdown", 6: "not_written"}

EXPORTS = [
    "cpm_solve_opts_default", "cpm_last_error", "cpm_ctx_create", "cpm_ctx_destroy", "cpm_ctx_synchronize",
    "cpm_ctx_device_bytes", "cpm_op_create_csr", "cpm_op_destroy", "cpm_op_rows", "cpm_op_format", "cpm_op_tile_stats", "cpm_op_gather_kind", "cpm_op_set_mn",
    "cpm_op_set_precond", "cpm_multiply", "cpm_multiply_device", "cpm_solve", "cpm_solve_device",
    "cpm_solve_batch", "cpm_solve_batch_device", "cpm_solve_stream_device", "cpm_device_alloc", "cpm_device_free", "cpm_memcpy_h2d",
    "cpm_memcpy_d2h", "cpm_tube_create", "cpm_tube_create_mesh", "cpm_tube_destroy", "cpm_tube_get_info", "cpm_tube_export",
    "cpm_tube_export_E", "cpm_tube_export_L", "cpm_tube_interp_matrix", "cpm_op_create_from_tube", "cpm_profile_egather", "cpm_profile_egather_groups", "cpm_profile_copy", "cpm_heat_step_device", "cpm_rd_params_default", "cpm_rd_step_device", "cpm_problem_create", "cpm_problem_destroy",
    "cpm_problem_get_info", "cpm_problem_rhs_points", "cpm_problem_subset", "cpm_problem_export",
    "cpm_op_create_from_problem", "cpm_geodesic_distance", "cpm_problem_plot_matrix", "cpm_obj_read",
    "cpm_polyline_chain", "cpm_tube_create_obj", "cpm_server_create", "cpm_server_destroy", "cpm_server_submit",
    "cpm_server_run", "cpm_problem_curve_rows", "cpm_problem_create_on", "cpm_problem_subset_size", "cpm_problem_ibc_parameter", "cpm_solve_device_noalloc",
]


class SolveOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iters", ctypes.c_int), ("precision", ctypes.c_int),
                ("update_delta", ctypes.c_double), ("reserved", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iters", ctypes.c_int), ("restarts", ctypes.c_int),
                ("updates", ctypes.c_int), ("rel_residual", ctypes.c_double), ("fallback_iters", ctypes.c_int)]

    def as_dict(self):
        return dict(status=STATUS.get(self.status, self.status), iters=self.iters, restarts=self.restarts,
                    updates=self.updates, rel_residual=self.rel_residual, fallback_iters=self.fallback_iters)


class RdParams(ctypes.Structure):
    _fields_ = [("F", ctypes.c_double), ("k", ctypes.c_double), ("nu_u", ctypes.c_double), ("nu_v", ctypes.c_double),
                ("dt", ctypes.c_double), ("method", ctypes.c_int), ("reserved", ctypes.c_int)]


class TubeInfo(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int64), ("n", ctypes.c_int64), ("N1d", ctypes.c_int64), ("nnz_L", ctypes.c_int64),
                ("levels", ctypes.c_int64), ("xstart", ctypes.c_double), ("dx", ctypes.c_double),
                ("tube_radius", ctypes.c_double), ("tube_radius_factor", ctypes.c_double),
                ("build_ms", ctypes.c_double), ("dim", ctypes.c_int), ("p", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IbcDesc(ctypes.Structure):
    _fields_ = [("type", ctypes.c_char_p), ("params", ctypes.c_void_p), ("nparams", ctypes.c_int),
                ("verts", ctypes.c_void_p), ("nverts", ctypes.c_int64), ("boundary_order", ctypes.c_int),
                ("boundary_type", ctypes.c_int * 2), ("on_ibc_tolerance", ctypes.c_double),
                ("dof_radius_scale", ctypes.c_double)]


class ProblemInfo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("tube_nodes", ctypes.c_int64), ("ibc_dofs", ctypes.c_int64),
                ("dirichlet_rows", ctypes.c_int64), ("identity_rows_start", ctypes.c_int64),
                ("tagging_size0", ctypes.c_int64), ("dof_size0", ctypes.c_int64), ("n_ibc", ctypes.c_int),
                ("build_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cpm error {code}: {msg}")
        self.code = code


_lib = None
vp = ctypes.c_void_p


def load(path=LIB_PATH):
    """Load libcpmgpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libcpmgpu.so not built at {path}; run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    L.cpm_last_error.restype = ctypes.c_char_p
    L.cpm_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.cpm_ctx_destroy.argtypes = [vp]
    L.cpm_ctx_synchronize.argtypes = [vp]
    L.cpm_ctx_device_bytes.argtypes = [vp]
    L.cpm_ctx_device_bytes.restype = ctypes.c_int64
    L.cpm_op_create_csr.argtypes = [vp, ctypes.c_int64, ctypes.c_int, vp, vp, vp, vp, vp, vp, vp,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(vp)]
    L.cpm_op_destroy.argtypes = [vp]
    L.cpm_op_rows.argtypes = [vp]
    L.cpm_op_rows.restype = ctypes.c_int64
    L.cpm_op_format.argtypes = [vp]
    L.cpm_op_format.restype = ctypes.c_int
    L.cpm_op_tile_stats.argtypes = [vp, vp]
    L.cpm_op_gather_kind.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cpm_op_gather_kind.restype = ctypes.c_int
    L.cpm_op_set_mn.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.cpm_op_set_precond.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.cpm_multiply.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double]
    L.cpm_multiply_device.argtypes = [vp, vp, vp, ctypes.c_double, ctypes.c_double]
    L.cpm_solve.argtypes = [vp, vp, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_solve_device.argtypes = [vp, vp, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_solve_device_noalloc.argtypes = [vp, vp, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_solve_batch.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_solve_batch_device.argtypes = [vp, ctypes.c_int, vp, vp, vp, ctypes.POINTER(SolveOpts),
                                         ctypes.POINTER(SolveInfo)]
    L.cpm_solve_opts_default.argtypes = [ctypes.POINTER(SolveOpts)]
    L.cpm_device_alloc.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(vp)]
    L.cpm_device_free.argtypes = [vp, vp]
    L.cpm_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_int64]
    L.cpm_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_int64]
    L.cpm_tube_create.argtypes = [vp, ctypes.c_char_p, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
    L.cpm_tube_create_mesh.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
    L.cpm_tube_create_obj.argtypes = [vp, ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      ctypes.c_double, ctypes.POINTER(vp)]
    L.cpm_obj_read.argtypes = [ctypes.c_char_p, vp, ctypes.POINTER(ctypes.c_int64), vp,
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)]
    L.cpm_polyline_chain.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, ctypes.POINTER(ctypes.c_int64)]
    L.cpm_tube_destroy.argtypes = [vp]
    L.cpm_server_create.argtypes = [vp, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(SolveOpts), ctypes.POINTER(vp)]
    L.cpm_server_destroy.argtypes = [vp]
    L.cpm_server_submit.argtypes = [vp, vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int64]
    L.cpm_server_run.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(SolveInfo), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.cpm_tube_get_info.argtypes = [vp, ctypes.POINTER(TubeInfo)]
    L.cpm_tube_export.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.cpm_tube_export_E.argtypes = [vp, vp, vp, vp]
    L.cpm_tube_export_L.argtypes = [vp, vp, vp, vp]
    L.cpm_tube_interp_matrix.argtypes = [vp, vp, ctypes.c_int64, vp, vp, vp]
    L.cpm_op_create_from_tube.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(vp)]
    L.cpm_heat_step_device.argtypes = [vp, ctypes.c_int, ctypes.c_double, vp, vp, ctypes.POINTER(SolveOpts),
                                       ctypes.POINTER(SolveInfo)]
    L.cpm_rd_params_default.restype = None
    L.cpm_rd_params_default.argtypes = [ctypes.c_double, ctypes.POINTER(RdParams)]
    L.cpm_rd_step_device.argtypes = [vp, ctypes.POINTER(RdParams), ctypes.c_int, ctypes.POINTER(vp),
                                     ctypes.POINTER(vp), vp, vp, ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_problem_plot_matrix.argtypes = [vp, vp, ctypes.c_int64, vp, vp, vp]
    L.cpm_geodesic_distance.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(SolveOpts), vp,
                                        ctypes.POINTER(SolveInfo)]
    L.cpm_solve_stream_device.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int,
                                          ctypes.POINTER(SolveOpts), ctypes.POINTER(SolveInfo)]
    L.cpm_profile_copy.argtypes = [vp, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.cpm_profile_egather.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double)]
    L.cpm_profile_egather_groups.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.cpm_problem_create.argtypes = [vp, ctypes.c_int, vp, ctypes.POINTER(vp)]
    L.cpm_problem_create_on.argtypes = [vp, vp, ctypes.c_int, vp, ctypes.POINTER(vp)]
    L.cpm_problem_destroy.argtypes = [vp]
    L.cpm_problem_curve_rows.argtypes = [vp, ctypes.c_int, vp]
    L.cpm_problem_get_info.argtypes = [vp, ctypes.POINTER(ProblemInfo)]
    L.cpm_problem_rhs_points.argtypes = [vp, vp, vp]
    L.cpm_problem_subset.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]
    L.cpm_problem_subset_size.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.cpm_problem_ibc_parameter.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int64, vp, vp]
    L.cpm_problem_export.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64)]
    L.cpm_op_create_from_problem.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(vp)]
    _lib = L
    return L


def _chk(rc):
    if rc != 0:
        raise CpmError(rc, load().cpm_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(vp)


def rd_params(dx, method=0, **kw):
    """Gray-Scott parameters with the reference's defaults at grid spacing dx; keyword overrides."""
    p = RdParams()
    load().cpm_rd_params_default(dx, ctypes.byref(p))
    p.method = method
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def opts(tol=1e-10, max_iters=100000, precision=CPM_FP64, update_delta=1e-2):
    o = SolveOpts()
    load().cpm_solve_opts_default(ctypes.byref(o))
    o.tol, o.max_iters, o.precision, o.update_delta, o.reserved = tol, max_iters, precision, update_delta, 0
    return o


def read_obj(path):
    """Wavefront OBJ as the reference reads it (SimplePolygonMesh::readMeshFromObjFile,
    src/SimplePolygonMesh.cpp:139-213): dict(verts (nv x 3), faces (nf x face_size, 0-based),
    kind "Triangulation" (face_size 3) or "Polyline" (2-vertex edges))."""
    L = load()
    nv, nf, fs = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int(0)
    bp = os.fsencode(path)
    _chk(L.cpm_obj_read(bp, None, ctypes.byref(nv), None, ctypes.byref(nf), ctypes.byref(fs)))
    verts = np.zeros((nv.value, 3))
    faces = np.zeros((nf.value, fs.value), np.int32)
    _chk(L.cpm_obj_read(bp, _ptr(verts), ctypes.byref(nv), _ptr(faces), ctypes.byref(nf), ctypes.byref(fs)))
    return dict(verts=verts, faces=faces, kind="Triangulation" if fs.value == 3 else "Polyline")


def polyline_chain(verts, edges):
    """Consecutive polyline edges -> the vertex chain "Polyline" surfaces and IBCs take; a closed
    polyline (src/SurfaceSpecifier.cpp:334-341) repeats its first vertex at the end."""
    L = load()
    vv = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    ee = np.ascontiguousarray(edges, np.int32).reshape(-1, 2)
    n = ctypes.c_int64(0)
    _chk(L.cpm_polyline_chain(_ptr(vv), len(vv), _ptr(ee), len(ee), None, ctypes.byref(n)))
    out = np.zeros((n.value, 3))
    _chk(L.cpm_polyline_chain(_ptr(vv), len(vv), _ptr(ee), len(ee), _ptr(out), ctypes.byref(n)))
    return out


class Context:
    def __init__(self, device=0):
        self._h = vp()
        _chk(load().cpm_ctx_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            load().cpm_ctx_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _chk(load().cpm_ctx_synchronize(self._h))

    def device_bytes(self):
        return load().cpm_ctx_device_bytes(self._h)

    def alloc(self, nbytes):
        p = vp()
        _chk(load().cpm_device_alloc(self._h, int(nbytes), ctypes.byref(p)))
        return p

    def free(self, p):
        _chk(load().cpm_device_free(self._h, p))

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        _chk(load().cpm_memcpy_h2d(self._h, dptr, _ptr(arr), arr.nbytes))

    def d2h(self, arr, dptr):
        _chk(load().cpm_memcpy_d2h(self._h, _ptr(arr), dptr, arr.nbytes))
        return arr

    def profile_copy(self, nbytes=1 << 28, reps=20):
        """Average ms of a plain device copy of nbytes (kernel k_copy16)."""
        ms = ctypes.c_double()
        _chk(load().cpm_profile_copy(self._h, int(nbytes), reps, ctypes.byref(ms)))
        return ms.value

    def solve_batch(self, ops, rhss, o=None):
        n = len(ops)
        o = o or opts()
        rhss = [np.ascontiguousarray(r, np.float64) for r in rhss]
        xs = [np.zeros_like(r) for r in rhss]
        OA = (vp * n)(*[op._h for op in ops])
        RA = (vp * n)(*[r.ctypes.data for r in rhss])
        XA = (vp * n)(*[x.ctypes.data for x in xs])
        infos = (SolveInfo * n)()
        _chk(load().cpm_solve_batch(self._h, n, OA, RA, XA, ctypes.byref(o), infos))
        return xs, [infos[i].as_dict() for i in range(n)]

    def solve_stream_device(self, op, d_rhs, d_x, lanes=8, o=None):
        """Continuous batching of many right-hand sides of one operator (FP32+refine)."""
        n = len(d_rhs)
        o = o or opts(precision=CPM_FP32_REFINE)
        RA = (vp * n)(*d_rhs)
        XA = (vp * n)(*d_x)
        infos = (SolveInfo * n)()
        _chk(load().cpm_solve_stream_device(op._h, n, RA, XA, lanes, ctypes.byref(o), infos))
        return [infos[i].as_dict() for i in range(n)]

    def solve_batch_device(self, ops, d_rhs, d_x, o=None):
        n = len(ops)
        o = o or opts()
        OA = (vp * n)(*[op._h for op in ops])
        RA = (vp * n)(*[p.value if isinstance(p, vp) else p for p in d_rhs])
        XA = (vp * n)(*[p.value if isinstance(p, vp) else p for p in d_x])
        infos = (SolveInfo * n)()
        _chk(load().cpm_solve_batch_device(self._h, n, OA, RA, XA, ctypes.byref(o), infos))
        return [infos[i].as_dict() for i in range(n)]


class Server:
    """Continuous batching across operators (FP32 Krylov + FP64 refinement, or FP64 with
    o.precision = CPM_FP64): requests (an operator and 1..lanes device right-hand sides) join the
    running iteration as pool memory allows and leave when they stop; results are bit-identical
    alone and in any mix."""

    def __init__(self, ctx, pool_bytes, lanes=4, o=None):
        self.ctx, self.lanes = ctx, lanes
        self._h = vp()
        self._keep = {}
        o = o or opts(precision=CPM_FP32_REFINE)
        _chk(load().cpm_server_create(ctx._h, int(pool_bytes), lanes, ctypes.byref(o), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            load().cpm_server_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, op, d_rhs, d_x, tag):
        """True if admitted; False if the pool has no room yet (submit again after run())."""
        n = len(d_rhs)
        RA = (vp * n)(*d_rhs)
        XA = (vp * n)(*d_x)
        rc = load().cpm_server_submit(self._h, op._h, n, RA, XA, int(tag))
        if rc == 1:
            return False
        _chk(rc)
        self._keep[int(tag)] = (op, RA, XA)
        return True

    def run(self, iterations=16, max_done=64):
        """Runs `iterations` iterations; returns ([(tag, [info dicts of its right-hand sides])], n_active)."""
        tags = (ctypes.c_int64 * max_done)()
        infos = (SolveInfo * (max_done * self.lanes))()
        nd, na = ctypes.c_int(0), ctypes.c_int(0)
        _chk(load().cpm_server_run(self._h, iterations, max_done, tags, infos, ctypes.byref(nd), ctypes.byref(na)))
        out = []
        for k in range(nd.value):
            t = int(tags[k])
            op, RA, _ = self._keep.pop(t)
            out.append((t, [infos[k * self.lanes + j].as_dict() for j in range(len(RA))]))
        return out, na.value


class Operator:
    """GPU operator A(m, n) = m I + n (Ldiag + Loff E) with identity rows (cpm::Solver::Init)."""

    def __init__(self, ctx, L, E, ident, m, n, nD=0, precond=PRECOND_DIAGONAL, precond_iters=0, handle=None):
        self.ctx = ctx
        if handle is not None:
            self._h = handle
            self.D = load().cpm_op_rows(handle)
            return
        D = len(L.ptr) - 1
        self._keep = [np.ascontiguousarray(a) for a in (L.ptr, L.col, L.val, E.ptr, E.col, E.val)]
        self._keep = [self._keep[0].astype(np.int32), self._keep[1].astype(np.int32), self._keep[2].astype(np.float64),
                      self._keep[3].astype(np.int32), self._keep[4].astype(np.int32), self._keep[5].astype(np.float64)]
        idn = np.ascontiguousarray(ident, np.uint8)
        self._h = vp()
        k = self._keep
        _chk(load().cpm_op_create_csr(ctx._h, D, nD, _ptr(k[0]), _ptr(k[1]), _ptr(k[2]), _ptr(k[3]), _ptr(k[4]),
                                      _ptr(k[5]), _ptr(idn), m, n, precond, precond_iters, ctypes.byref(self._h)))
        self._keep = None
        self.D = D

    def close(self):
        if self._h:
            load().cpm_op_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def format(self):
        """'ell' (general sparse rows) or 'tp' (tensor-product stencils of a tube-built operator)."""
        return {0: "ell", 1: "tp"}[load().cpm_op_format(self._h)]

    def tile_info(self):
        """LDS tiles of the E gather (tensor-product operators)."""
        v = np.zeros(6, np.int64)
        _chk(load().cpm_op_tile_stats(self._h, _ptr(v)))
        keys = ("shared_stencil_tiles", "items", "shared_stencil_items", "list_rows", "row_tiles", "row_list_rows")
        return {k: int(x) for k, x in zip(keys, v)}

    def gather_kernel(self, precision=CPM_FP64, nrhs=1):
        """Name of the E-gather kernel of single-operator launches of this shape."""
        k = load().cpm_op_gather_kind(self._h, precision, nrhs)
        if k < 0:
            raise CpmError(f"no E gather for precision {precision}, {nrhs} right-hand sides")
        t = "float" if precision == CPM_FP32_REFINE else "double"
        return f"{'k_egather_grp' if k == 2 else 'k_egather'}<{t},{nrhs}>"

    def set_mn(self, m, n):
        _chk(load().cpm_op_set_mn(self._h, m, n))

    def set_precond(self, precond, iters=0):
        _chk(load().cpm_op_set_precond(self._h, precond, iters))

    def multiply(self, x, m, n):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        _chk(load().cpm_multiply(self._h, _ptr(x), _ptr(y), m, n))
        return y

    def multiply_device(self, dx, dy, m, n):
        _chk(load().cpm_multiply_device(self._h, dx, dy, m, n))

    def solve(self, rhs, o=None):
        rhs = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(rhs)
        info = SolveInfo()
        _chk(load().cpm_solve(self._h, _ptr(rhs), _ptr(x), ctypes.byref(o or opts()), ctypes.byref(info)))
        return x, info.as_dict()

    def profile_egather(self, precision=CPM_FP64, nrhs=1, reps=50, groups=1):
        """Average E-gather launch (ms) and its algorithmic bytes, `groups` copies of the operator per
        launch (the streamed solves of 16 and 32 lanes run 2 and 4)."""
        ms = ctypes.c_double()
        by = ctypes.c_double()
        _chk(load().cpm_profile_egather_groups(self._h, precision, nrhs, groups, reps, ctypes.byref(ms),
                                               ctypes.byref(by)))
        return ms.value, by.value

    def heat_step_device(self, method, dt, d_g, d_u, o=None):
        """One HeatWithIBC time step on device vectors (0 explicit, 1 backward Euler, 2 Crank-Nicolson)."""
        info = SolveInfo()
        _chk(load().cpm_heat_step_device(self._h, method, dt, d_g, d_u, ctypes.byref(o or opts()), ctypes.byref(info)))
        return info.as_dict()

    def rd_step_device(self, params, d_us, d_vs, d_gu, d_gv, o=None):
        """One Gray-Scott reaction-diffusion step of every (u, v) field pair on the device
        (params: RdParams, method 0 explicit as the reference, 1 implicit diffusion)."""
        n = len(d_us)
        us, vs = (ctypes.c_void_p * n)(*d_us), (ctypes.c_void_p * n)(*d_vs)
        infos = (SolveInfo * (2 * n))()
        _chk(load().cpm_rd_step_device(self._h, ctypes.byref(params), n, us, vs, d_gu, d_gv, ctypes.byref(o or opts()),
                                       infos))
        return [i.as_dict() for i in infos]

    def solve_device(self, d_rhs, d_x, o=None):
        info = SolveInfo()
        _chk(load().cpm_solve_device(self._h, d_rhs, d_x, ctypes.byref(o or opts()), ctypes.byref(info)))
        return info.as_dict()

    def solve_device_noalloc(self, d_rhs, d_x, o=None):
        """solve_device with its workspaces in the library's fixed device scratch (no allocation)."""
        info = SolveInfo()
        _chk(load().cpm_solve_device_noalloc(self._h, d_rhs, d_x, ctypes.byref(o or opts()), ctypes.byref(info)))
        return info.as_dict()


class CSR:
    def __init__(self, ptr, col, val):
        self.ptr, self.col, self.val = ptr, col, val

    @property
    def rows(self):
        return len(self.ptr) - 1

    def matvec(self, x):
        rows = np.repeat(np.arange(self.rows), np.diff(self.ptr))
        return np.bincount(rows, weights=self.val * x[self.col], minlength=self.rows)


class Tube:
    """GPU-built closest-point tube (Tube(SurfaceSpecifier&, p, dx))."""

    def __init__(self, ctx, surface="Sphere", dx=0.1, params=None, bbox=None, p=3, verts=None, faces=None,
                 obj=None):
        """surface "Triangulation" takes verts (nv x 3) and faces (nf x 3); obj, a path, takes the
        surface an OBJ file holds (triangles or a polyline; SurfaceSpecifier(file)); bbox None then
        gives the reference's box [min(0, coordinates), max(0, coordinates)]."""
        self.ctx = ctx
        self._h = vp()
        if obj is not None:
            bb = bbox if bbox is not None else (0.0, 0.0)
            _chk(load().cpm_tube_create_obj(ctx._h, os.fsencode(obj), float(bb[0]), float(bb[1]), p, dx,
                                            ctypes.byref(self._h)))
        elif surface == "Triangulation":
            vv = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
            ff = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
            bb = bbox if bbox is not None else (0.0, 0.0)
            _chk(load().cpm_tube_create_mesh(ctx._h, _ptr(vv), len(vv), _ptr(ff), len(ff), float(bb[0]), float(bb[1]),
                                             p, dx, ctypes.byref(self._h)))
        else:
            par = np.ascontiguousarray(params if params is not None else [0.0], np.float64)
            npar = 0 if params is None else len(par)
            bb = bbox if bbox is not None else (0.0, 0.0)
            vv = np.ascontiguousarray(verts if verts is not None else np.zeros((1, 3)), np.float64)
            nv = 0 if verts is None else len(vv)
            _chk(load().cpm_tube_create(ctx._h, surface.encode(), _ptr(par), npar, float(bb[0]), float(bb[1]), 3,
                                        _ptr(vv), nv, p, dx, ctypes.byref(self._h)))
        inf = TubeInfo()
        _chk(load().cpm_tube_get_info(self._h, ctypes.byref(inf)))
        self.info = inf.as_dict()
        self.N = self.info["nodes"]

    def close(self):
        if self._h:
            load().cpm_tube_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        N = self.N
        rec = np.zeros(N, np.int64)
        x = np.zeros((N, 3))
        cp = np.zeros((N, 3))
        dist = np.zeros(N)
        bdy = np.zeros(N, np.int32)
        nbr = np.zeros((N, 6), np.int32)
        bpt = np.zeros(N, np.int32)
        _chk(load().cpm_tube_export(self._h, _ptr(rec), _ptr(x), _ptr(cp), _ptr(dist), _ptr(bdy), _ptr(nbr), _ptr(bpt)))
        return dict(rec=rec, x=x, cp=cp, dist=dist, bdy=bdy, nbr=nbr, bpt=bpt)

    def E(self):
        W = (self.info["p"] + 1) ** 3
        ptr = np.zeros(self.N + 1, np.int32)
        col = np.zeros(self.N * W, np.int32)
        val = np.zeros(self.N * W)
        _chk(load().cpm_tube_export_E(self._h, _ptr(ptr), _ptr(col), _ptr(val)))
        return CSR(ptr, col, val)

    def L(self):
        nnz = self.info["nnz_L"]
        ptr = np.zeros(self.N + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        _chk(load().cpm_tube_export_L(self._h, _ptr(ptr), _ptr(col), _ptr(val)))
        return CSR(ptr, col, val)

    def interp(self, xq):
        xq = np.ascontiguousarray(xq, np.float64).reshape(-1, 3)
        W = (self.info["p"] + 1) ** 3
        nq = len(xq)
        ptr = np.zeros(nq + 1, np.int32)
        col = np.zeros(nq * W, np.int32)
        val = np.zeros(nq * W)
        _chk(load().cpm_tube_interp_matrix(self._h, _ptr(xq), nq, _ptr(ptr), _ptr(col), _ptr(val)))
        return CSR(ptr, col, val)

    def operator(self, m, n, precond=PRECOND_DIAGONAL, precond_iters=0):
        h = vp()
        _chk(load().cpm_op_create_from_tube(self._h, m, n, precond, precond_iters, ctypes.byref(h)))
        return Operator(self.ctx, None, None, None, m, n, handle=h)


class IBC:
    """An interior boundary curve (SurfaceSpecifier + IBCMetaData)."""

    @classmethod
    def from_obj(cls, path, **meta):
        """Polyline curve of an OBJ file ("l" records, SurfaceSpecifier(file))."""
        d = read_obj(path)
        if d["kind"] != "Polyline":
            raise CpmError(-1, f"{path}: an interior boundary curve needs polyline edges, not triangles")
        return cls("Polyline", verts=polyline_chain(d["verts"], d["faces"]), **meta)

    def __init__(self, type, params=None, verts=None, boundary_order=2, boundary_type=(0, 0), on_ibc_tolerance=0.0,
                 dof_radius_scale=1.0):
        self.type, self.params, self.verts = type, params, verts
        self.boundary_order, self.boundary_type = boundary_order, boundary_type
        self.on_ibc_tolerance, self.dof_radius_scale = on_ibc_tolerance, dof_radius_scale


class Problem:
    """PoissonWithIBC on a GPU tube: subsets, directions and the IBC-modified operator built on the device."""

    def __init__(self, tube, ibcs, ctx=None):
        """ctx: build on this context (same device as the tube; e.g. a build thread's) instead of the
        tube's; the problem's operators then belong to it."""
        self.tube = tube
        n = len(ibcs)
        descs = (IbcDesc * n)()
        self._keep = []
        for d, i in zip(descs, ibcs):
            par = np.ascontiguousarray(i.params if i.params is not None else [0.0], np.float64)
            vv = np.ascontiguousarray(i.verts if i.verts is not None else np.zeros((1, 3)), np.float64)
            self._keep += [par, vv]
            d.type = i.type.encode()
            d.params = par.ctypes.data
            d.nparams = 0 if i.params is None else len(par)
            d.verts = vv.ctypes.data
            d.nverts = 0 if i.verts is None else len(vv)
            d.boundary_order = i.boundary_order
            d.boundary_type[0], d.boundary_type[1] = i.boundary_type
            d.on_ibc_tolerance = i.on_ibc_tolerance
            d.dof_radius_scale = i.dof_radius_scale
        self._h = vp()
        self.ctx = ctx or tube.ctx
        _chk(load().cpm_problem_create_on(ctx._h if ctx is not None else None, tube._h, n, ctypes.cast(descs, vp),
                                          ctypes.byref(self._h)))
        inf = ProblemInfo()
        _chk(load().cpm_problem_get_info(self._h, ctypes.byref(inf)))
        self.info = inf.as_dict()
        self.D = self.info["rows"]

    def close(self):
        if self._h:
            load().cpm_problem_destroy(self._h)
            self._h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rhs_points(self):
        pts = np.zeros((self.D, 3))
        kind = np.zeros(self.D, np.int32)
        _chk(load().cpm_problem_rhs_points(self._h, _ptr(pts), _ptr(kind)))
        return pts, kind

    def curve_rows(self, c):
        """Rows of curve c: (first ghost row, ghost rows, first Dirichlet-value row, Dirichlet-value rows)."""
        r = np.zeros(4, np.int64)
        _chk(load().cpm_problem_curve_rows(self._h, c, _ptr(r)))
        return tuple(int(v) for v in r)

    def ibc_parameter(self, c, pts, vertex_param=None):
        """Surface::IBCParameter of points on curve c (GPU): normalised arc length, or with vertex_param
        (one value per polyline vertex) the nearest vertex's value."""
        P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros(len(P))
        vpp = None
        if vertex_param is not None:
            vpa = np.ascontiguousarray(vertex_param, np.float64)
            vpp = _ptr(vpa)
        _chk(load().cpm_problem_ibc_parameter(self._h, c, _ptr(P), len(P), vpp, _ptr(out)))
        return out

    def subset(self, c=0, which=0):
        """Tagging (which 0) or DOF (1) subset of curve c: tube indices, closest points on the curve, directions."""
        nn = ctypes.c_int64()
        _chk(load().cpm_problem_subset_size(self._h, c, which, ctypes.byref(nn)))
        n = nn.value
        s = np.zeros(n, np.int32)
        cp = np.zeros((n, 3))
        d = np.zeros((n, 3))
        _chk(load().cpm_problem_subset(self._h, c, which, _ptr(s), _ptr(cp), _ptr(d)))
        return s, cp, d

    def export(self):
        nl = ctypes.c_int64()
        ne = ctypes.c_int64()
        _chk(load().cpm_problem_export(self._h, None, None, None, None, None, None, None, ctypes.byref(nl),
                                       ctypes.byref(ne)))
        D = self.D
        Lp, Lc, Lv = np.zeros(D + 1, np.int32), np.zeros(nl.value, np.int32), np.zeros(nl.value)
        Ep, Ec, Ev = np.zeros(D + 1, np.int32), np.zeros(ne.value, np.int32), np.zeros(ne.value)
        idn = np.zeros(D, np.uint8)
        _chk(load().cpm_problem_export(self._h, _ptr(Lp), _ptr(Lc), _ptr(Lv), _ptr(Ep), _ptr(Ec), _ptr(Ev), _ptr(idn),
                                       ctypes.byref(nl), ctypes.byref(ne)))
        return CSR(Lp, Lc, Lv), CSR(Ep, Ec, Ev), idn

    def operator(self, m, n, precond=PRECOND_DIAGONAL, precond_iters=0):
        h = vp()
        _chk(load().cpm_op_create_from_problem(self._h, m, n, precond, precond_iters, ctypes.byref(h)))
        return Operator(self.ctx, None, None, None, m, n, handle=h)

    def plot_matrix(self, xq):
        """Plotting rows at points xq (n x 3) with the curves' ghost columns (AddIBCPlotInterpolation)."""
        xq = np.ascontiguousarray(xq, np.float64).reshape(-1, 3)
        W = (self.tube.info["p"] + 1) ** 3
        ptr = np.zeros(len(xq) + 1, np.int32)
        col = np.zeros(len(xq) * W, np.int32)
        val = np.zeros(len(xq) * W)
        _chk(load().cpm_problem_plot_matrix(self._h, _ptr(xq), len(xq), _ptr(ptr), _ptr(col), _ptr(val)))
        return CSR(ptr, col, val)

    def geodesic(self, dt, sweeps=2000, o=None):
        """GeodesicDistance::ComputeDistance before its plotting interpolation: phi on the D rows
        (host array) and the infos of the heat solve and the three Poisson solves."""
        ctx = self.tube.ctx
        d_phi = ctx.alloc(8 * self.D)
        infos = (SolveInfo * 4)()
        try:
            _chk(load().cpm_geodesic_distance(self._h, dt, sweeps, ctypes.byref(o or opts()), d_phi, infos))
            phi = ctx.d2h(np.zeros(self.D), d_phi)
        finally:
            ctx.free(d_phi)
        return phi, [i.as_dict() for i in infos]
