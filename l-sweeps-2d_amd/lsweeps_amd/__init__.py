"""Python binding of the MI355X L-sweeps library (ctypes over include/lsweeps_hip.h).

The HIP library ``liblsweeps_hip.so`` must have been built in-tree
(``__graft_entry__.build()`` or ``python l-sweeps-2d_amd/lsweeps_amd/build.py``).
There is no CPU fallback: every call goes to the HIP library, and importing
this module raises if the library is missing.
"""
import ctypes as C
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblsweeps_hip.so")
if not os.path.exists(_LIB_PATH):
    raise ImportError(f"HIP library not built: {_LIB_PATH} (run __graft_entry__.build())")
_lib = C.CDLL(_LIB_PATH)

LSW_PREC_FP64 = 0
LSW_PREC_FP32 = 1
LSW_ORDER_OLDEST = 0
LSW_ORDER_BATCH = 1
ERRORS = {1: "ARG", 2: "HIP", 3: "STATE", 4: "NUMERIC", 5: "CAPACITY", 6: "UNSUPPORTED", 7: "INTERNAL"}

EXPORTED = [
    "lsw_last_error", "lsw_version", "lsw_device_count", "lsw_device_memory",
    "lsw_lu_create", "lsw_lu_set_csc_z", "lsw_lu_factorize", "lsw_lu_solve_z", "lsw_lu_solve_dev", "lsw_lu_solve_dev_noalloc",
    "lsw_lu_info",
    "lsw_lu_destroy",
    "lsw_dd_create", "lsw_dd_create_config", "lsw_dd_create_host", "lsw_dd_destroy", "lsw_dd_factorize", "lsw_dd_sizes",
    "lsw_dd_global_coords", "lsw_dd_cell_info", "lsw_dd_assemble_rhs", "lsw_dd_assemble_rhs_dev",
    "lsw_dd_apply_precond", "lsw_dd_apply_precond_dev", "lsw_dd_apply_a", "lsw_dd_apply_a_dev",
    "lsw_dd_stats", "lsw_dd_sync", "lsw_dd_profile_precond", "lsw_dd_profile_precond_block",
    "lsw_dd_apply_precond_block", "lsw_dd_apply_precond_block_dev",
    "lsw_gmres_solve", "lsw_gmres_solve_dev",
    "lsw_config_num_problems", "lsw_config_problem", "lsw_problem_run",
    "lsw_gmres_solve_block", "lsw_gmres_solve_block_dev", "lsw_dd_assemble_rhs_block_dev",
    "lsw_server_create", "lsw_server_destroy", "lsw_server_add_problem", "lsw_server_submit",
    "lsw_server_wait", "lsw_server_get_stats", "lsw_server_problem_bytes", "lsw_dd_write_vtk",
    "lsw_dd_set_graphs", "lsw_dd_autotune",
]


class LswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{ERRORS.get(code, code)}] {msg}")
        self.code = code


class DDParams(C.Structure):
    _fields_ = [
        ("domain", C.c_double * 4),
        ("ncells_x", C.c_int), ("ncells_y", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int),
        ("omega", C.c_double),
        ("min_wavelengths_in_pml", C.c_double),
        ("order", C.c_int),
        ("precision", C.c_int),
        ("medium_kind", C.c_int),
        ("velocity", C.POINTER(C.c_float)),
        ("vel_rows", C.c_int), ("vel_cols", C.c_int),
        ("vel_domain", C.c_double * 4),
        ("geometry", C.c_int),
    ]


class GmresParams(C.Structure):
    _fields_ = [("mode", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
                ("restart", C.c_int), ("verbose", C.c_int)]


class GmresResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("precond_applies", C.c_int), ("matvecs", C.c_int),
                ("estimate", C.c_double), ("true_residual", C.c_double)]


class ProblemSpec(C.Structure):
    _fields_ = [("dd", DDParams), ("csv_medium", C.c_int), ("solver_mode", C.c_int), ("tol", C.c_double),
                ("max_iter", C.c_int), ("restart", C.c_int), ("verbose", C.c_int), ("nsources", C.c_int)]


class ProblemResult(C.Structure):
    _fields_ = [("setup_seconds", C.c_double), ("factorize_seconds", C.c_double),
                ("gmres_seconds", C.c_double), ("iterations", C.c_int), ("estimate", C.c_double),
                ("true_residual", C.c_double), ("dofs", C.c_longlong), ("Gx", C.c_int), ("Gy", C.c_int)]


class ServerOptions(C.Structure):
    _fields_ = [("budget_bytes", C.c_double), ("slots", C.c_int), ("restart", C.c_int), ("tol", C.c_double),
                ("max_iter", C.c_int), ("precision", C.c_int), ("workers", C.c_int), ("order", C.c_int)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int), ("problem", C.c_int), ("nsrc", C.c_int), ("max_iterations", C.c_int),
                ("max_true_residual", C.c_double), ("submit_s", C.c_double), ("start_s", C.c_double),
                ("end_s", C.c_double), ("message", C.c_char * 160)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_longlong), ("completed", C.c_longlong), ("solves", C.c_longlong),
                ("loads", C.c_longlong), ("evictions", C.c_longlong), ("steps", C.c_longlong),
                ("reserved_bytes", C.c_double), ("peak_reserved_bytes", C.c_double), ("resident", C.c_int),
                ("merged_steps", C.c_longlong), ("max_active", C.c_int), ("setup_s", C.c_double),
                ("factor_s", C.c_double), ("block_s", C.c_double), ("step_s", C.c_double)]


_lib.lsw_last_error.restype = C.c_char_p
_lib.lsw_version.restype = C.c_char_p
_P = C.c_void_p
_D = C.POINTER(C.c_double)


def _sig(name, *args):
    f = getattr(_lib, name)
    f.argtypes = list(args)
    f.restype = C.c_int


_sig("lsw_device_count", C.POINTER(C.c_int))
_sig("lsw_device_memory", C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int)
_sig("lsw_lu_create", C.c_int, C.POINTER(_P))
_sig("lsw_lu_set_csc_z", _P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _D)
_sig("lsw_lu_factorize", _P)
_sig("lsw_lu_solve_z", _P, C.c_int, _D, _D)
_sig("lsw_lu_solve_dev", _P, C.c_int, _P, _P, _P)
_sig("lsw_lu_solve_dev_noalloc", _P, C.c_int, _P, _P, _P)
_sig("lsw_lu_info", _P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong),
     C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("lsw_lu_destroy", _P)
_sig("lsw_dd_create", C.POINTER(DDParams), C.POINTER(_P))
_sig("lsw_dd_create_host", C.POINTER(DDParams), C.POINTER(_P))
_sig("lsw_dd_create_config", C.c_char_p, C.c_int, C.POINTER(_P))
_sig("lsw_dd_destroy", _P)
_sig("lsw_dd_factorize", _P)
_sig("lsw_dd_sizes", _P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _D)
_sig("lsw_dd_global_coords", _P, _D, _D)
_sig("lsw_dd_cell_info", _P, C.c_int, C.POINTER(C.c_int))
_sig("lsw_dd_assemble_rhs", _P, C.c_int, _D, _D)
_sig("lsw_dd_assemble_rhs_dev", _P, C.c_int, _D, _P, _P)
_sig("lsw_dd_apply_precond", _P, _D, _D)
_sig("lsw_dd_apply_precond_dev", _P, _P, _P, _P)
_sig("lsw_dd_apply_a", _P, _D, _D)
_sig("lsw_dd_apply_a_dev", _P, _P, _P, _P)
_sig("lsw_dd_stats", _P, C.POINTER(C.c_longlong), _D, C.POINTER(C.c_longlong))
_sig("lsw_dd_sync", _P)
_sig("lsw_dd_set_graphs", _P, C.c_int)
_sig("lsw_dd_autotune", _P, C.c_int, _D)
_sig("lsw_dd_profile_precond", _P, _P, _P, _D)
_sig("lsw_dd_profile_precond_block", _P, C.c_int, _P, _P, _D)
_sig("lsw_dd_apply_precond_block", _P, C.c_int, _D, _D)
_sig("lsw_dd_apply_precond_block_dev", _P, C.c_int, _P, _P, _P)
_sig("lsw_gmres_solve", _P, C.POINTER(GmresParams), _D, _D, C.POINTER(GmresResult), _D, C.c_int)
_sig("lsw_gmres_solve_dev", _P, C.POINTER(GmresParams), _P, _P, _P, C.POINTER(GmresResult))


_sig("lsw_gmres_solve_block", _P, C.POINTER(GmresParams), C.c_int, _D, _D, C.POINTER(GmresResult))
_sig("lsw_gmres_solve_block_dev", _P, C.POINTER(GmresParams), C.c_int, _P, _P, _P, C.POINTER(GmresResult))
_sig("lsw_dd_assemble_rhs_block_dev", _P, C.c_int, _D, _P, _P)
_sig("lsw_config_num_problems", C.c_char_p, C.POINTER(C.c_int))
_sig("lsw_config_problem", C.c_char_p, C.c_int, C.POINTER(ProblemSpec), _D, C.c_int)
_sig("lsw_problem_run", C.c_char_p, C.c_int, C.c_int, _D, C.c_longlong, C.POINTER(ProblemResult))
_sig("lsw_dd_write_vtk", _P, _D, C.c_char_p)
_sig("lsw_server_create", C.POINTER(ServerOptions), C.POINTER(_P))
_sig("lsw_server_destroy", _P)
_sig("lsw_server_add_problem", _P, C.POINTER(DDParams), C.POINTER(C.c_int), C.POINTER(C.c_int),
     C.POINTER(C.c_int))
_sig("lsw_server_submit", _P, C.c_int, C.c_int, _D, _D, C.POINTER(C.c_longlong))
_sig("lsw_server_wait", _P, C.c_longlong, C.POINTER(RequestResult))
_sig("lsw_server_problem_bytes", _P, C.c_int, _D, _D)
_sig("lsw_server_get_stats", _P, C.POINTER(ServerStats))


def config_num_problems(path):
    n = C.c_int()
    _check(_lib.lsw_config_num_problems(path.encode(), C.byref(n)))
    return n.value


def config_problem(path, pid=0, max_sources=4096):
    """Parse problem `pid` of a libconfig file (host only)."""
    spec = ProblemSpec()
    src = np.zeros(2 * max_sources)
    _check(_lib.lsw_config_problem(path.encode(), pid, C.byref(spec), _dptr(src), max_sources))
    d = spec.dd
    return dict(domain=((d.domain[0], d.domain[1]), (d.domain[2], d.domain[3])),
                ncells=(d.ncells_x, d.ncells_y), nels=(d.nx, d.ny), omega=d.omega,
                min_wl_in_pml=d.min_wavelengths_in_pml, order=d.order, precision=d.precision,
                csv_medium=bool(spec.csv_medium), solver_mode=spec.solver_mode, tol=spec.tol,
                max_iter=spec.max_iter, restart=spec.restart, verbose=bool(spec.verbose),
                sources=src[:2 * min(spec.nsources, max_sources)].reshape(-1, 2).copy())


def problem_run(path, pid=0, verbose=False, want_solution=True):
    """Problem::read + Problem::solve of a config file on the GPU."""
    r = ProblemResult()
    x = None
    if want_solution:
        spec = ProblemSpec()
        _check(_lib.lsw_config_problem(path.encode(), pid, C.byref(spec), None, 0))
        n = 0
        if not spec.csv_medium:  # geometry only depends on the PML width, known without the file
            g = DDCheckerboard(spec.dd, device=False)
            n = g.N
        x = np.zeros(max(n, 1) if n else 1 << 26, dtype=np.complex128)
    _check(_lib.lsw_problem_run(path.encode(), pid, int(verbose),
                                _dptr(x.view(np.float64)) if x is not None else None,
                                len(x) if x is not None else 0, C.byref(r)))
    out = dict(setup_s=r.setup_seconds, factorize_s=r.factorize_seconds, gmres_s=r.gmres_seconds,
               iterations=r.iterations, estimate=r.estimate, true_residual=r.true_residual, dofs=r.dofs,
               Gx=r.Gx, Gy=r.Gy)
    if x is not None:
        out["x"] = x[:r.dofs].copy()
    return out


def _check(rc):
    if rc != 0:
        raise LswError(rc, _lib.lsw_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(_D)


def lib():
    return _lib


def device_count():
    n = C.c_int(0)
    _check(_lib.lsw_device_count(C.byref(n)))
    return n.value


# --------------------------------------------------------------------------
def device_memory(release=False):
    """(live, cached) device bytes the library holds; release=True first
    hands the cached blocks back to the HIP runtime."""
    live, cached = C.c_longlong(), C.c_longlong()
    _check(_lib.lsw_device_memory(C.byref(live), C.byref(cached), int(release)))
    return live.value, cached.value


class LocalLinearSolverHIP:
    """setMatrix / factorize / solve of one 5-point structured-grid matrix."""

    def __init__(self, precision=LSW_PREC_FP64):
        self._h = _P()
        _check(_lib.lsw_lu_create(precision, C.byref(self._h)))

    def set_matrix(self, A):
        import scipy.sparse as sp
        A = sp.csc_matrix(A)
        A.sort_indices()
        n = A.shape[0]
        self.n = n
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
        rowind = np.ascontiguousarray(A.indices, dtype=np.int32)
        vals = np.ascontiguousarray(A.data, dtype=np.complex128).view(np.float64)
        _check(_lib.lsw_lu_set_csc_z(self._h, n, int(A.nnz), colptr.ctypes.data_as(C.POINTER(C.c_int)),
                                     rowind.ctypes.data_as(C.POINTER(C.c_int)), _dptr(vals)))

    def factorize(self):
        _check(_lib.lsw_lu_factorize(self._h))

    def solve(self, b):
        b = np.asarray(b, dtype=np.complex128)
        one = b.ndim == 1
        # n x nrhs column-major == C-contiguous (nrhs, n)
        Bt = np.ascontiguousarray(b.reshape(self.n, -1).T)
        Xt = np.zeros_like(Bt)
        _check(_lib.lsw_lu_solve_z(self._h, Bt.shape[0], _dptr(Bt.view(np.float64)),
                                   _dptr(Xt.view(np.float64))))
        return Xt[0] if one else Xt.T

    def solve_dev(self, nrhs, b_ptr, x_ptr, stream=0):
        """Device pointers: n x nrhs column-major complex128 (b_ptr == x_ptr allowed)."""
        _check(_lib.lsw_lu_solve_dev(self._h, int(nrhs), b_ptr, x_ptr, stream or None))

    def solve_dev_noalloc(self, nrhs, b_ptr, x_ptr, stream=0):
        """solve_dev with no device allocation (the library's fixed scratch array)."""
        _check(_lib.lsw_lu_solve_dev_noalloc(self._h, int(nrhs), b_ptr, x_ptr, stream or None))

    def info(self):
        nx, ny, nodes, h = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        fe = C.c_longlong()
        _check(_lib.lsw_lu_info(self._h, C.byref(nx), C.byref(ny), C.byref(fe), C.byref(nodes), C.byref(h)))
        return dict(nx=nx.value, ny=ny.value, factor_entries=fe.value, nodes=nodes.value, height=h.value)

    def __del__(self):
        # (at interpreter exit the module's globals may already be cleared)
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lsw_lu_destroy(self._h)
            self._h = None


# --------------------------------------------------------------------------
def dd_params(domain=((0.0, 1.0), (0.0, 1.0)), ncells=(4, 4), nels=(9, 9), omega=2 * np.pi * 4.1,
              min_wl_in_pml=2.0, precision=LSW_PREC_FP64, velocity=None, vel_domain=None, order=2,
              general=False):
    p = DDParams()
    p.geometry = 1 if general else 0
    p.domain[:] = [domain[0][0], domain[0][1], domain[1][0], domain[1][1]]
    p.ncells_x, p.ncells_y = int(ncells[0]), int(ncells[1])
    p.nx, p.ny = int(nels[0]), int(nels[1])
    p.omega = float(omega)
    p.min_wavelengths_in_pml = float(min_wl_in_pml)
    p.order = order
    p.precision = precision
    if velocity is None:
        p.medium_kind = 0
        p._vel_keep = None
    else:
        v = np.ascontiguousarray(velocity, dtype=np.float32)
        p.medium_kind = 1
        p._vel_keep = v
        p.velocity = v.ctypes.data_as(C.POINTER(C.c_float))
        p.vel_rows, p.vel_cols = v.shape
        vd = vel_domain if vel_domain is not None else domain
        p.vel_domain[:] = [vd[0][0], vd[0][1], vd[1][0], vd[1][1]]
    return p


class DDCheckerboard:
    """A checkerboard DD with every cell resident on the GPU."""

    def __init__(self, params, device=True, config=None):
        """params: DDParams; or config = (input file path, problem id) for the
        DD a reference-format input file describes (params is then None)."""
        self.params = params
        self._h = _P()
        if config is not None:
            _check(_lib.lsw_dd_create_config(str(config[0]).encode(), int(config[1]), C.byref(self._h)))
        else:
            fn = _lib.lsw_dd_create if device else _lib.lsw_dd_create_host
            _check(fn(C.byref(params), C.byref(self._h)))
        gx, gy, nc = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        _check(_lib.lsw_dd_sizes(self._h, C.byref(gx), C.byref(gy), C.byref(nc), C.byref(ms)))
        self.Gx, self.Gy, self.ncells, self.mesh_size = gx.value, gy.value, nc.value, ms.value
        self.N = self.Gx * self.Gy

    @property
    def handle(self):
        return self._h

    def coords(self):
        xs = np.zeros(self.Gx)
        ys = np.zeros(self.Gy)
        _check(_lib.lsw_dd_global_coords(self._h, _dptr(xs), _dptr(ys)))
        return xs, ys

    def set_graphs(self, on=True):
        _check(_lib.lsw_dd_set_graphs(self._h, int(bool(on))))

    def autotune(self, nrhs=1):
        """Time M^{-1} for blocks of nrhs right-hand sides under alternative
        launch shapes and keep the fastest for that block width."""
        out = np.zeros(11)
        _check(_lib.lsw_dd_autotune(self._h, int(nrhs), _dptr(out)))
        return {"default_ms": float(out[0]), "best_ms": float(out[1]), "tried": int(out[2]),
                "latency_waves": int(out[3]), "wide_budget": int(out[4]), "wide_cap": int(out[5]),
                "mfma": bool(out[6]), "latency_rows": int(out[7]), "mfma_wide_cols": int(out[8]),
                "mfma_pair": bool(out[9]), "mfma_group": bool(out[10])}

    def write_vtk(self, x, path):
        """<path>.vtk (OutputVTK layout) and <path>.lswgrid (full precision)."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        assert x.size == self.N
        _check(_lib.lsw_dd_write_vtk(self._h, _dptr(x.view(np.float64)), str(path).encode()))

    def cell_info(self, c):
        info = (C.c_int * 38)()
        _check(_lib.lsw_dd_cell_info(self._h, c, info))
        a = list(info)
        return dict(I=a[0], J=a[1], nx=a[2], ny=a[3], ox=a[4], oy=a[5],
                    vlr=[a[6:14], a[14:22]], vud=[a[22:30], a[30:38]])

    def factorize(self):
        _check(_lib.lsw_dd_factorize(self._h))

    def assemble_rhs(self, sources):
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 2))
        b = np.zeros(self.N, dtype=np.complex128)
        _check(_lib.lsw_dd_assemble_rhs(self._h, src.shape[0], _dptr(src), _dptr(b.view(np.float64))))
        return b

    def apply_precond(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        _check(_lib.lsw_dd_apply_precond(self._h, _dptr(x.view(np.float64)), _dptr(y.view(np.float64))))
        return y

    def apply_precond_block(self, X):
        """X: (nrhs, N) complex array; returns M^{-1} applied to every row."""
        X = np.ascontiguousarray(X, dtype=np.complex128)
        Y = np.zeros_like(X)
        _check(_lib.lsw_dd_apply_precond_block(self._h, X.shape[0], _dptr(X.view(np.float64)),
                                               _dptr(Y.view(np.float64))))
        return Y

    def apply_precond_block_dev(self, nrhs, x_ptr, y_ptr, stream=0):
        _check(_lib.lsw_dd_apply_precond_block_dev(self._h, nrhs, x_ptr, y_ptr, stream or None))

    def apply_a(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros_like(x)
        _check(_lib.lsw_dd_apply_a(self._h, _dptr(x.view(np.float64)), _dptr(y.view(np.float64))))
        return y

    def stats(self):
        s, fb = C.c_longlong(), C.c_longlong()
        b = C.c_double()
        _check(_lib.lsw_dd_stats(self._h, C.byref(s), C.byref(b), C.byref(fb)))
        return dict(solves_per_precond=s.value, solve_bytes_per_precond=b.value, factor_bytes=fb.value)

    def gmres(self, b, mode=0, tol=1e-8, max_iter=30, restart=30, verbose=False):
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        p = GmresParams(mode, tol, max_iter, restart, int(verbose))
        r = GmresResult()
        log = np.zeros(max_iter + 2 + (max_iter // max(restart, 1) + 2) * 2)
        _check(_lib.lsw_gmres_solve(self._h, C.byref(p), _dptr(b.view(np.float64)),
                                    _dptr(x.view(np.float64)), C.byref(r), _dptr(log), len(log)))
        n = r.iterations + 1 if mode == 0 else len(log)
        return x, dict(iterations=r.iterations, precond_applies=r.precond_applies, matvecs=r.matvecs,
                       estimate=r.estimate, true_residual=r.true_residual, log=log[:n].copy())

    def gmres_block(self, B, tol=1e-6, max_iter=60, restart=30, verbose=False):
        """Flexible GMRES for the rows of B (nrhs, N) with shared block preconditioning."""
        B = np.ascontiguousarray(B, dtype=np.complex128)
        X = np.zeros_like(B)
        p = GmresParams(1, tol, max_iter, restart, int(verbose))
        res = (GmresResult * B.shape[0])()
        _check(_lib.lsw_gmres_solve_block(self._h, C.byref(p), B.shape[0], _dptr(B.view(np.float64)),
                                          _dptr(X.view(np.float64)), res))
        return X, [dict(iterations=r.iterations, precond_applies=r.precond_applies, matvecs=r.matvecs,
                        estimate=r.estimate, true_residual=r.true_residual) for r in res]

    def gmres_block_dev(self, nrhs, b_ptr, x_ptr, tol=1e-6, max_iter=60, restart=30, stream=0):
        p = GmresParams(1, tol, max_iter, restart, 0)
        res = (GmresResult * nrhs)()
        _check(_lib.lsw_gmres_solve_block_dev(self._h, C.byref(p), nrhs, b_ptr, x_ptr, stream or None, res))
        return [dict(iterations=r.iterations, precond_applies=r.precond_applies, matvecs=r.matvecs,
                     estimate=r.estimate, true_residual=r.true_residual) for r in res]

    def assemble_rhs_block_dev(self, sources, b_ptr, stream=0):
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 2))
        _check(_lib.lsw_dd_assemble_rhs_block_dev(self._h, src.shape[0], _dptr(src), b_ptr, stream or None))

    # device-pointer variants (pointers are integers, e.g. torch tensor.data_ptr())
    def apply_precond_dev(self, x_ptr, y_ptr, stream=0):
        _check(_lib.lsw_dd_apply_precond_dev(self._h, x_ptr, y_ptr, stream or None))

    def apply_a_dev(self, x_ptr, y_ptr, stream=0):
        _check(_lib.lsw_dd_apply_a_dev(self._h, x_ptr, y_ptr, stream or None))

    def assemble_rhs_dev(self, sources, b_ptr, stream=0):
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 2))
        _check(_lib.lsw_dd_assemble_rhs_dev(self._h, src.shape[0], _dptr(src), b_ptr, stream or None))

    def gmres_dev(self, b_ptr, x_ptr, mode=0, tol=1e-8, max_iter=30, restart=30, stream=0):
        p = GmresParams(mode, tol, max_iter, restart, 0)
        r = GmresResult()
        _check(_lib.lsw_gmres_solve_dev(self._h, C.byref(p), b_ptr, x_ptr, stream or None, C.byref(r)))
        return dict(iterations=r.iterations, precond_applies=r.precond_applies, matvecs=r.matvecs,
                    estimate=r.estimate, true_residual=r.true_residual)

    def sync(self):
        _check(_lib.lsw_dd_sync(self._h))

    def profile_precond_dev(self, x_ptr, y_ptr, nrhs=1):
        """Per-launch HIP-event timing of the batched local solves of one M^{-1}
        (x_ptr, y_ptr: nrhs consecutive global vectors on the device)."""
        out = np.zeros(9)
        if nrhs == 1:
            _check(_lib.lsw_dd_profile_precond(self._h, x_ptr, y_ptr, _dptr(out)))
        else:
            _check(_lib.lsw_dd_profile_precond_block(self._h, int(nrhs), x_ptr, y_ptr, _dptr(out)))
        return dict(fwd_ms=out[0], fwd_launches=int(out[1]), fwd_bytes=out[2],
                    bwd_ms=out[3], bwd_launches=int(out[4]), bwd_bytes=out[5],
                    fwd_flops=out[6], bwd_flops=out[7], ideal_ms=out[8])

    def __del__(self):
        # (at interpreter exit the module's globals may already be cleared)
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lsw_dd_destroy(self._h)
            self._h = None


# --------------------------------------------------------------------------
class Server:
    """Continuous-batching solve server (lsw_server_*): a device factor cache
    keyed by registered problem, admission by device bytes, and per-Arnoldi-step
    batching of every right-hand side of a problem.  submit() returns at once;
    wait() blocks.  Solutions land in host arrays that the server owns until
    the request is waited for."""

    def __init__(self, budget_bytes=250e9, slots=16, restart=16, tol=1e-6, max_iter=200,
                 precision=LSW_PREC_FP32, workers=4, batch_order=False):
        o = ServerOptions(float(budget_bytes), int(slots), int(restart), float(tol), int(max_iter),
                          int(precision), int(workers), LSW_ORDER_BATCH if batch_order else LSW_ORDER_OLDEST)
        self._h = _P()
        _check(_lib.lsw_server_create(C.byref(o), C.byref(self._h)))
        self._grids = {}
        self._out = {}

    def add_problem(self, params):
        pid, gx, gy = C.c_int(), C.c_int(), C.c_int()
        _check(_lib.lsw_server_add_problem(self._h, C.byref(params), C.byref(pid), C.byref(gx), C.byref(gy)))
        self._grids[pid.value] = (gx.value, gy.value)
        return pid.value

    def grid(self, pid):
        return self._grids[pid]

    def problem_bytes(self, pid):
        st, pk = C.c_double(), C.c_double()
        _check(_lib.lsw_server_problem_bytes(self._h, int(pid), C.byref(st), C.byref(pk)))
        return st.value, pk.value

    def submit(self, pid, sources, want_solution=True):
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 2))
        x = None
        if want_solution:
            gx, gy = self._grids[pid]
            x = np.zeros((len(src), gx * gy), dtype=np.complex128)
        rid = C.c_longlong()
        _check(_lib.lsw_server_submit(self._h, int(pid), len(src), _dptr(src.reshape(-1)),
                                      _dptr(x.view(np.float64)) if x is not None else None, C.byref(rid)))
        self._out[rid.value] = (x, src)
        return rid.value

    def wait(self, rid):
        r = RequestResult()
        _check(_lib.lsw_server_wait(self._h, C.c_longlong(rid), C.byref(r)))
        x, _ = self._out.pop(rid)
        return dict(status=r.status, message=r.message.decode(), problem=r.problem, nsrc=r.nsrc,
                    max_iterations=r.max_iterations, max_true_residual=r.max_true_residual,
                    submit_s=r.submit_s, start_s=r.start_s, end_s=r.end_s, latency_s=r.end_s - r.submit_s,
                    x=x)

    def stats(self):
        st = ServerStats()
        _check(_lib.lsw_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            _check(_lib.lsw_server_destroy(self._h))
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
