"""ctypes view of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product library never
loads it.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_set_mesh.restype = None
        L.orc_set_mesh.argtypes = [f64p, ctypes.c_int64, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS"),
                                   ctypes.c_int64]
        L.orc_mesh_cp.argtypes = [f64p, ctypes.c_int64, f64p, f64p]
        L.orc_build_ibc_plot_E.argtypes = [ctypes.c_void_p, f64p, ctypes.c_int64]
        L.orc_geodesic.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, f64p,
                                   np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        L.orc_problem_poisson.restype = ctypes.c_void_p
        L.orc_problem_poisson.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_int, ctypes.c_int]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_info.argtypes = [ctypes.c_void_p, i64p, f64p]
        L.orc_tube.argtypes = [ctypes.c_void_p, i64p, f64p, f64p, f64p, i64p, i64p, i64p]
        L.orc_csr.argtypes = [ctypes.c_void_p, ctypes.c_int, i32p, i32p, f64p]
        L.orc_ident.argtypes = [ctypes.c_void_p, u8p]
        L.orc_build_explicit.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_build_plot_E.argtypes = [ctypes.c_void_p]
        L.orc_multiply.argtypes = [ctypes.c_int64, ctypes.c_int, i32p, i32p, f64p, i32p, i32p, f64p, u8p,
                                   ctypes.c_double, ctypes.c_double, f64p, f64p]
        L.orc_solve.argtypes = [ctypes.c_int64, ctypes.c_int, i32p, i32p, f64p, i32p, i32p, f64p, u8p,
                                ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, f64p, f64p,
                                ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_double)]
        L.orc_poisson_problem_error.restype = ctypes.c_double
        L.orc_poisson_problem_error.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
        L.orc_weights_1d.argtypes = [f64p, ctypes.c_double, ctypes.c_int, f64p]
        L.orc_problem_poisson_ibc.restype = ctypes.c_void_p
        L.orc_problem_poisson_ibc.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                              ctypes.c_int]
        L.orc_ibc_info.argtypes = [ctypes.c_void_p, i64p]
        L.orc_ibc_subset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, i64p, f64p, f64p]
        L.orc_ibc_parameter.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.orc_ibc_subset_size.restype = ctypes.c_int64
        L.orc_ibc_subset_size.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_rhs_points.argtypes = [ctypes.c_void_p, f64p, np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")]
        L.orc_heat_run.restype = ctypes.c_int64
        L.orc_heat_run.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.orc_rd_run.restype = ctypes.c_int64
        L.orc_rd_run.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, f64p, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_void_p]
        L.orc_rd_run_problem.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.orc_heat_study_order.restype = ctypes.c_double
        L.orc_heat_study_order.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_heat_problem_error.restype = ctypes.c_double
        L.orc_heat_problem_error.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_ibc_study_order.restype = ctypes.c_double
        L.orc_ibc_study_order.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        _LIB = L
    return _LIB


def _err():
    return lib().orc_last_error().decode()


@dataclass
class CSR:
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray

    @property
    def rows(self):
        return len(self.ptr) - 1

    def to_dense(self, ncols=None):
        n = ncols or self.rows
        A = np.zeros((self.rows, n))
        for r in range(self.rows):
            s, e = self.ptr[r], self.ptr[r + 1]
            A[r, self.col[s:e]] += self.val[s:e]
        return A

    def matvec(self, x):
        rows = np.repeat(np.arange(self.rows), np.diff(self.ptr))
        return np.bincount(rows, weights=self.val * x[self.col], minlength=self.rows)


class Problem:
    """Poisson problem (problems/Poisson.cpp) built by the oracle: tube, Lfull, Efull, identity rows."""

    def __init__(self, surface="Sphere", dx=0.1, params=None, bbox=None, dim=3, verts=None,
                 bdy_order=0, which_bdy=(), p=3):
        L = lib()
        use_def = params is None and verts is None
        par = np.ascontiguousarray(params if params is not None else [0.0], dtype=np.float64)
        bb = bbox if bbox is not None else (0.0, 0.0)
        vv = np.ascontiguousarray(verts if verts is not None else np.zeros((1, dim)), dtype=np.float64)
        wb = np.ascontiguousarray(list(which_bdy) or [0], dtype=np.int64)
        self._h = L.orc_problem_poisson(surface.encode(), int(use_def), par.ctypes.data, len(par),
                                        float(bb[0]), float(bb[1]), dim, vv.ctypes.data, len(vv), dx,
                                        bdy_order, wb.ctypes.data, len(which_bdy), p)
        if not self._h:
            raise RuntimeError(_err())
        self._info()

    def _info(self):
        ints = np.zeros(9, np.int64)
        dbl = np.zeros(4)
        lib().orc_info(self._h, ints, dbl)
        (self.N, self.D, self.nnzL, self.nnzE, self.n, self.N1d, self.dim, self.p, self.nnz_extra) = map(int, ints)
        self.xstart, self.r_t, self.rfac, self.dx = map(float, dbl)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def tube(self):
        N, d = self.N, self.dim
        rec = np.zeros(N, np.int64)
        x = np.zeros(N * d)
        cp = np.zeros(N * d)
        dist = np.zeros(N)
        bdy = np.zeros(N, np.int64)
        nbr = np.zeros(N * 6, np.int64)
        bpt = np.zeros(N, np.int64)
        lib().orc_tube(self._h, rec, x, cp, dist, bdy, nbr, bpt)
        return dict(rec=rec, x=x.reshape(N, d), cp=cp.reshape(N, d), dist=dist, bdy=bdy,
                    nbr=nbr.reshape(N, 6), bpt=bpt)

    def _csr(self, which, rows, nnz):
        ptr = np.zeros(rows + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1))
        lib().orc_csr(self._h, which, ptr, col, val)
        return CSR(ptr, col[:nnz], val[:nnz])

    def L(self):
        return self._csr(0, self.D, self.nnzL)

    def E(self):
        return self._csr(1, self.D, self.nnzE)

    def ident(self):
        out = np.zeros(self.D, np.uint8)
        lib().orc_ident(self._h, out)
        return out

    def explicit_operator(self, m, n, nD=0):
        nnz = lib().orc_build_explicit(self._h, m, n, nD)
        if nnz < 0:
            raise RuntimeError(_err())
        self._info()
        return self._csr(2, self.D, nnz)

    def plot_E(self):
        rows = lib().orc_build_plot_E(self._h)
        if rows < 0:
            raise RuntimeError(_err())
        self._info()
        return self._csr(2, rows, self.nnz_extra)


def set_mesh(verts, faces):
    """Triangle mesh (nv x 3 vertices, nf x 3 indices) for the next "Triangulation" surface on this thread."""
    v = np.ascontiguousarray(verts, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    lib().orc_set_mesh(v.ravel(), len(v), f.ravel(), len(f))


def mesh_cp(verts, faces, x):
    """Closest points and distances on a triangle mesh (brute force over triangles, Ericson 5.1.5)."""
    set_mesh(verts, faces)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    cp = np.zeros_like(x)
    d = np.zeros(len(x))
    if lib().orc_mesh_cp(x.ravel(), len(x), cp.ravel(), d) < 0:
        raise RuntimeError(_err())
    return cp, d


def mesh_problem(verts, faces, dx, bbox=None, p=3):
    """Poisson problem on a triangle mesh (SurfaceSpecifier(".obj"), closest points by brute force)."""
    set_mesh(verts, faces)
    if bbox is None:
        return Problem("Triangulation", dx, p=p)
    return Problem("Triangulation", dx, params=[0.0], bbox=bbox, p=p)


def multiply(L, E, ident, x, m, n, nD=0):
    y = np.zeros_like(x)
    if lib().orc_multiply(len(x), nD, L.ptr, L.col, L.val, E.ptr, E.col, E.val,
                          np.ascontiguousarray(ident, np.uint8), m, n, np.ascontiguousarray(x), y) != 0:
        raise RuntimeError(_err())
    return y


def set_parallel(vec=False, apply=False, threads=0):
    """Host-thread parallelism of the solver (timing only; results are unchanged)."""
    lib().orc_set_parallel(int(vec), int(apply), int(threads))


def solve(L, E, ident, b, m, n, nD=0, precond=1, precond_iters=0, tol=1e-10, max_it=100000):
    x = np.zeros_like(b)
    it = ctypes.c_int(0)
    rr = ctypes.c_double(0)
    if lib().orc_solve(len(b), nD, L.ptr, L.col, L.val, E.ptr, E.col, E.val,
                       np.ascontiguousarray(ident, np.uint8), m, n, precond, precond_iters,
                       np.ascontiguousarray(b), x, tol, max_it, ctypes.byref(it), ctypes.byref(rr)) != 0:
        raise RuntimeError(_err())
    return x, it.value, rr.value


def poisson_problem_error(problem, dx):
    it = ctypes.c_int(0)
    e = lib().orc_poisson_problem_error(problem, dx, ctypes.byref(it))
    if e < 0:
        raise RuntimeError(_err())
    return e, it.value


def weights_1d(xs, xq):
    xs = np.ascontiguousarray(xs, np.float64)
    w = np.zeros(len(xs))
    lib().orc_weights_1d(xs, float(xq), len(xs), w)
    return w


class IBC:
    """One interior boundary curve: SurfaceSpecifier + IBCMetaData (src/IBCMetaData.h:9-24)."""

    def __init__(self, type, params=None, verts=None, boundary_order=2, boundary_type=(0, 0), on_ibc_tolerance=0.0):
        self.type, self.params, self.verts = type, params, verts
        self.boundary_order, self.boundary_type, self.on_ibc_tolerance = boundary_order, boundary_type, on_ibc_tolerance


class IBCProblem(Problem):
    """PoissonWithIBC (problems/PoissonWithIBC.cpp) built by the oracle."""

    def __init__(self, surface="Torus", dx=0.1, params=None, bbox=None, ibcs=(), p=3):
        L = lib()
        use_def = params is None
        par = np.ascontiguousarray(params if params is not None else [0.0], dtype=np.float64)
        bb = bbox if bbox is not None else (0.0, 0.0)
        types = (ctypes.c_char_p * len(ibcs))(*[i.type.encode() for i in ibcs])
        ip = np.ascontiguousarray(np.concatenate([np.asarray(i.params or [], float) for i in ibcs] + [np.zeros(1)]))
        nip = np.ascontiguousarray([len(i.params or []) for i in ibcs], np.int32)
        vv = [np.asarray(i.verts, float).reshape(-1, 3) for i in ibcs if i.verts is not None]
        iv = np.ascontiguousarray(np.concatenate(vv + [np.zeros((1, 3))]).ravel())
        nv = np.ascontiguousarray([len(i.verts) if i.verts is not None else 0 for i in ibcs], np.int64)
        bo = np.ascontiguousarray([i.boundary_order for i in ibcs], np.int32)
        bt = np.ascontiguousarray([t for i in ibcs for t in i.boundary_type], np.int32)
        tl = np.ascontiguousarray([i.on_ibc_tolerance for i in ibcs], np.float64)
        self._h = L.orc_problem_poisson_ibc(surface.encode(), int(use_def), par.ctypes.data, len(par), float(bb[0]),
                                            float(bb[1]), len(ibcs), types, ip.ctypes.data, nip.ctypes.data,
                                            iv.ctypes.data, nv.ctypes.data, bo.ctypes.data, bt.ctypes.data,
                                            tl.ctypes.data, dx, p)
        if not self._h:
            raise RuntimeError(_err())
        self._info()
        ii = np.zeros(6, np.int64)
        L.orc_ibc_info(self._h, ii)
        self.G, self.K, self.identity_rows_start, self.n_ibc, self.n_tag0, self.n_dof0 = map(int, ii)

    def subset(self, c=0, which=0):
        n = lib().orc_ibc_subset_size(self._h, c, which)
        if n < 0:
            raise ValueError(f"no curve {c}")
        s = np.zeros(n, np.int64)
        cp = np.zeros(n * 3)
        d = np.zeros(n * 3)
        lib().orc_ibc_subset(self._h, c, which, s, cp, d)
        return s, cp.reshape(-1, 3), d.reshape(-1, 3)

    def plot_matrix(self, xq):
        """Plotting rows at xq (n x 3) with the curves' ghost columns (AddIBCPlotInterpolation)."""
        xq = np.ascontiguousarray(xq, np.float64).reshape(-1, 3)
        rows = lib().orc_build_ibc_plot_E(self._h, xq.ravel(), len(xq))
        if rows < 0:
            raise RuntimeError(_err())
        self._info()
        return self._csr(2, rows, self.nnz_extra)

    def geodesic(self, dt, sweeps=2000):
        """GeodesicDistance::ComputeDistance before its plotting interpolation: phi on the D rows
        and the iteration counts of the heat solve and the three Poisson solves."""
        phi = np.zeros(self.D)
        it = np.zeros(4, np.int32)
        if lib().orc_geodesic(self._h, dt, sweeps, phi, it) < 0:
            raise RuntimeError(_err())
        return phi, it

    def ibc_parameter(self, c, pts, vertex_param=None):
        """Surface::IBCParameter (src/Surface.cpp:2731-2788) of points on curve c."""
        P = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros(len(P))
        vp = None if vertex_param is None else np.ascontiguousarray(vertex_param, np.float64)
        if lib().orc_ibc_parameter(self._h, c, P.ctypes.data, len(P), None if vp is None else vp.ctypes.data,
                                   0 if vp is None else len(vp), out.ctypes.data) < 0:
            raise RuntimeError(_err())
        return out

    def rd_run(self, method, u0, v0, F, k, nu_u, nu_v, dt, steps):
        """Gray-Scott steps on this problem from (u0, v0), Dirichlet values 0 (ReactionDiffusionWithIBC::
        TimeStep for method 0, implicit diffusion for method 1)."""
        prm = np.array([F, k, nu_u, nu_v, dt], np.float64)
        u, v = np.array(u0, np.float64), np.array(v0, np.float64)
        assert len(u) == self.D and len(v) == self.D
        if lib().orc_rd_run_problem(self._h, method, prm.ctypes.data, steps, u.ctypes.data, v.ctypes.data) < 0:
            raise RuntimeError(_err())
        return u, v

    def rhs_points(self):
        pts = np.zeros(self.D * 3)
        kind = np.zeros(self.D, np.int32)
        lib().orc_rhs_points(self._h, pts, kind)
        return pts.reshape(-1, 3), kind


def heat_study_order(method, boundary_order=2, boundary_type=0):
    """HeatProblemsWithIBC problems 3-5 (method 0 explicit, 1 backward Euler, 2 Crank-Nicolson) with a
    Dirichlet (type 0) or Neumann (type 1) curve of boundary order 1 or 2."""
    o = lib().orc_heat_study_order(method, boundary_order, boundary_type)
    if o < 0:
        raise RuntimeError(_err())
    return o


def heat_run(method, dx, steps, boundary_order=2, boundary_type=0):
    """Final state after `steps` steps of the heat study's system at dx (and the step size)."""
    dt = ctypes.c_double(0)
    D = lib().orc_heat_run(method, dx, 0, boundary_order, boundary_type, None, ctypes.byref(dt))
    if D < 0:
        raise RuntimeError(_err())
    u = np.zeros(D)
    if lib().orc_heat_run(method, dx, steps, boundary_order, boundary_type, u.ctypes.data, ctypes.byref(dt)) < 0:
        raise RuntimeError(_err())
    return u, dt.value


def heat_problem_error(method, dx, boundary_order=2, boundary_type=0):
    """Max error at the plot points against exp(-2 t) cos(phi) at t = 0.1."""
    e = lib().orc_heat_problem_error(method, dx, boundary_order, boundary_type)
    if e < 0:
        raise RuntimeError(_err())
    return e


def ibc_study_order(problem, g, boundary_order=2):
    """PoissonProblemsWithIBC problem 2 (sphere + circle), 5 (sphere + point) or 8 (sphere + arc) on
    the seeded geometry g, Dirichlet curves of boundary order 1 or 2."""
    it = ctypes.c_int(0)
    o = lib().orc_ibc_study_order(problem, g, boundary_order, ctypes.byref(it))
    if o < 0:
        raise RuntimeError(_err())
    return o


def rd_run(method, dx, u0, v0, F, k, nu_u, nu_v, dt, steps, boundary_order=2, boundary_type=0):
    """Gray-Scott reaction-diffusion steps on the heat study's system (sphere and
    Circle3D curve) from (u0, v0); method 0 explicit, 1 implicit diffusion."""
    prm = np.array([F, k, nu_u, nu_v, dt], np.float64)
    D = lib().orc_rd_run(method, dx, boundary_order, boundary_type, prm, 0, None, None)
    if D < 0:
        raise RuntimeError(_err())
    assert len(u0) == D and len(v0) == D
    u, v = np.array(u0, np.float64), np.array(v0, np.float64)
    if lib().orc_rd_run(method, dx, boundary_order, boundary_type, prm, steps, u.ctypes.data, v.ctypes.data) < 0:
        raise RuntimeError(_err())
    return u, v
