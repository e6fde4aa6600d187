"""MI355X-native nodal elasticity solve (gfx950 HIP kernels behind a C ABI).

This module binds ``libela.so`` (built in-tree by ``anisotropicfracture_amd.build``)
with ctypes and mirrors the reference's operator interface for Python callers:

* :class:`Elastic` -- ``Operator::Elastic<T>`` (src/Operator/Elastic.H:20-171)
  with the node-operator base ``Operator<Grid::Node>`` (src/Operator/Operator.H:30-128):
  ``SetModel``, ``SetBC``, ``SetUniform``, ``prepareForSolve``, ``Fapply``,
  ``Diagonal``, ``Fsmooth``, ``restriction``, ``interpolation``, ``residual``.
* :class:`Linear` -- ``Solver::Nonlocal::Linear`` (src/Solver/Nonlocal/Linear.H:19-143).
* :class:`Batch` / :class:`RVE` -- many independent box problems solved
  together (one launch per cycle kernel over all members); each member is the
  same problem and iteration as ``Linear.solve`` on its own operator.

Every computation runs on the GPU through the library; there is no CPU
fallback.  If the library is missing, importing works but the first use
raises ``RuntimeError``.

Host arrays are numpy float64 of shape ``(ncomp, nz+1, ny+1, nx+1)`` in 3D or
``(ncomp, ny+1, nx+1)`` in 2D (component-major, x fastest).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libela.so")

ISO, MM, DIAG = 0, 1, 2
DISP, TRAC, PERIODIC, NEUMANN = 0, 1, 2, 3
F64, F32 = 0, 1
SOLVE_FP64, SOLVE_MIXED, ALLOC_BOTH = 0, 1, 2
CYCLE_F, CYCLE_V = 0, 1
OPT_MARCH, OPT_KZ, OPT_XCD, OPT_MASK, OPT_GRAPH, OPT_FOLD, OPT_ROWS, OPT_PREFETCH, OPT_GJ = 0, 1, 2, 3, 4, 5, 6, 7, 8
OPT_GRAIN_DIAG, OPT_PROLONG = 9, 10

ERRORS = {1: "BADARG", 2: "OOM", 3: "HIP", 4: "NONCONVERGED", 5: "NAN", 6: "STATE"}


class ElaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ela error %s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class NonConverged(ElaError):
    pass


class _Domain(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n", ctypes.c_int * 3), ("dx", ctypes.c_double * 3)]


class _Layout(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n", ctypes.c_int * 3), ("nodes", ctypes.c_int * 3),
                ("pitch", ctypes.c_int64), ("plane", ctypes.c_int64), ("cstride", ctypes.c_int64),
                ("dx", ctypes.c_double * 3)]


class Stats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("converged", ctypes.c_int), ("mg_levels", ctypes.c_int),
                ("mode", ctypes.c_int), ("res0", ctypes.c_double), ("res", ctypes.c_double),
                ("seconds", ctypes.c_double), ("setup_seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(ctypes.Structure):
    _fields_ = [("arrival", ctypes.c_double), ("n", ctypes.c_int), ("nseeds", ctypes.c_int),
                ("seeds", ctypes.c_void_p), ("q", ctypes.c_int), ("R", ctypes.c_void_p),
                ("C11", ctypes.c_double), ("C12", ctypes.c_double), ("C44", ctypes.c_double),
                ("crack", ctypes.c_int), ("crack_p", ctypes.c_double * 3), ("crack_nrm", ctypes.c_double * 3),
                ("crack_tv", ctypes.c_double * 3), ("crack_width", ctypes.c_double), ("crack_g", ctypes.c_double * 3),
                ("G", ctypes.c_double * 9), ("u_out", ctypes.c_void_p), ("max_cycles", ctypes.c_int)]


class ServeSummary(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("groups", ctypes.c_int), ("admitted", ctypes.c_int),
                ("peak_active", ctypes.c_int), ("peak_bytes", ctypes.c_size_t), ("t_admit", ctypes.c_double),
                ("t_prepare", ctypes.c_double), ("t_step", ctypes.c_double), ("t_retire", ctypes.c_double),
                ("t_idle", ctypes.c_double)]


class PlotLevel(ctypes.Structure):
    _fields_ = [("domain", ctypes.c_int * 3), ("origin", ctypes.c_int * 3), ("nodes", ctypes.c_int * 3),
                ("data", ctypes.c_void_p), ("nboxes", ctypes.c_int), ("boxes", ctypes.c_void_p)]


class Result(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_double), ("finished", ctypes.c_double), ("cycles", ctypes.c_int),
                ("status", ctypes.c_int), ("res0", ctypes.c_double), ("res", ctypes.c_double)]


_lib = None

# exported symbol -> (argtypes, restype)
_P = ctypes.c_void_p
_I, _D, _SZ = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
SIGNATURES = {
    "ela_last_error": ([], ctypes.c_char_p),
    "ela_ctx_create": ([_I, _SZ, ctypes.POINTER(_P)], _I),
    "ela_ctx_destroy": ([_P], _I),
    "ela_ctx_bytes_in_use": ([_P], _SZ),
    "ela_hier_create": ([_P, ctypes.POINTER(_Domain), _I, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_hier_destroy": ([_P], _I),
    "ela_hier_levels": ([_P], _I),
    "ela_level_info": ([_P, _I, ctypes.POINTER(_Layout)], _I),
    "ela_hier_bytes": ([_P], _SZ),
    "ela_set_model_field": ([_P, _P], _I),
    "ela_set_model_voronoi_cubic": ([_P, _I, _P, _I, _P, _D, _D, _D], _I),
    "ela_set_model_degraded_iso": ([_P, _D, _D, _I, _P], _I),
    "ela_degrade_material": ([_P, _D, _D, _I, _P, _P], _I),
    "ela_combine_cubic": ([_P, _I, _P, _P, _P, _P, _P], _I),
    "ela_newton_w_dw": ([_P, _P, _P, _D, _P, _P, _P], _I),
    "ela_dev_alloc": ([_P, _SZ, ctypes.POINTER(_P)], _I),
    "ela_dev_free": ([_P, _P], _I),
    "ela_memcpy": ([_P, _P, _SZ], _I),
    "ela_set_bc": ([_P, _P, _P], _I),
    "ela_set_uniform": ([_P, _I], _I),
    "ela_prepare": ([_P], _I),
    "ela_field_alloc": ([_P, _I, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_field_free": ([_P, _P], _I),
    "ela_upload": ([_P, _I, _I, _I, _P, _P], _I),
    "ela_download": ([_P, _I, _I, _I, _P, _P], _I),
    "ela_field_zero": ([_P, _I, _I, _I, _P], _I),
    "ela_fapply": ([_P, _I, _I, _P, _P, _P], _I),
    "ela_diagonal": ([_P, _I, _I, _P, _P], _I),
    "ela_fsmooth": ([_P, _I, _I, _P, _P, _I, _D, _P], _I),
    "ela_restriction": ([_P, _I, _I, _P, _P, _P], _I),
    "ela_interpolation": ([_P, _I, _I, _P, _P, _P], _I),
    "ela_residual": ([_P, _I, _I, _P, _P, _P, ctypes.POINTER(_D), _P], _I),
    "ela_residual_noalloc": ([_P, _I, _I, _P, _P, _P, ctypes.POINTER(_D), _P], _I),
    "ela_bc_init": ([_P, _P, _I], _I),
    "ela_solve": ([_P, _P, _P, _D, _D, _I, _I, ctypes.POINTER(Stats)], _I),
    "ela_solve_device": ([_P, _P, _P, _D, _D, _I, _I, ctypes.POINTER(Stats)], _I),
    "ela_set_cycle": ([_P, _I, _I, _I], _I),
    "ela_set_timing": ([_P, _I], _I),
    "ela_set_krylov": ([_P, _I], _I),
    "ela_set_stiffness_mode": ([_P, _I], _I),
    "ela_set_option": ([_P, _I, _I], _I),
    "ela_autotune": ([_P, _I, _P], _I),
    "ela_batch_autotune": ([_P, _I, _P, _I, _P], _I),
    "ela_batch_graph_stats": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "ela_kernel_times": ([_P, ctypes.POINTER(_D), ctypes.POINTER(_I), ctypes.POINTER(_D), ctypes.POINTER(_I)], _I),
    "ela_time_fapply": ([_P, _I, _I, ctypes.POINTER(_D)], _I),
    "ela_probe_copy": ([_P, _SZ, _I, ctypes.POINTER(_D)], _I),
    "ela_probe_copy_ex": ([_P, _SZ, _I, _I, ctypes.POINTER(_D)], _I),
    "ela_probe_barrier": ([_P, _I, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)], _I),
    "ela_hier_create_ex": ([_P, ctypes.POINTER(_Domain), _I, _I, _I, _I, ctypes.POINTER(_P)], _I),
    # batched independent solves
    "ela_batch_create": ([_P, _I, ctypes.POINTER(_P)], _I),
    "ela_batch_destroy": ([_P], _I),
    "ela_batch_bytes_in_use": ([_P], _SZ),
    "ela_rve_bytes": ([_I, _I, _I, _I], _SZ),
    "ela_rve_create": ([_P, _I, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_rve_destroy": ([_P], _I),
    "ela_rve_bytes_in_use": ([_P], _SZ),
    "ela_rve_levels": ([_P], _I),
    "ela_rve_set_bc_types": ([_P, _P], _I),
    "ela_rve_set_uniform": ([_P, _I], _I),
    "ela_rve_set_voronoi_cubic": ([_P, _I, _P, _I, _P, _D, _D, _D], _I),
    "ela_rve_set_crack": ([_P, _P, _P, _P, _D, _D, _D, _D], _I),
    "ela_rve_set_rhs_affine": ([_P, _P], _I),
    "ela_rve_upload": ([_P, _I, _P], _I),
    "ela_rve_download": ([_P, _I, _P], _I),
    "ela_rve_grain_ids": ([_P, _P], _I),
    "ela_batch_prepare": ([_P, _I, _P], _I),
    "ela_batch_fapply": ([_P, _I, _P, _I], _I),
    "ela_batch_submit": ([_P, _P, _D, _D, _I], _I),
    "ela_batch_step": ([_P, ctypes.POINTER(_I)], _I),
    "ela_batch_solve": ([_P, _I, _P, _D, _D, _I], _I),
    "ela_rve_status": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(Stats)], _I),
    "ela_batch_set_timing": ([_P, _I], _I),
    "ela_batch_set_option": ([_P, _I, _I], _I),
    "ela_normalize": ([_P, _I, _I, _P, _P], _I),
    "ela_postprocess": ([_P, _I, _I, _P, _P, _P], _I),
    "ela_newton_rhs": ([_P, _P, _P, _P, _P, _P], _I),
    "ela_reflux": ([ctypes.POINTER(_Layout), _P, ctypes.POINTER(_Layout), _P, _P, _I, _I, _P], _I),
    "ela_average_down": ([ctypes.POINTER(_Layout), _P, ctypes.POINTER(_Layout), _P, _P, _I, _I, _P], _I),
    "ela_reflux_window": ([ctypes.POINTER(_Layout), ctypes.POINTER(_Layout), _P, _P, ctypes.POINTER(_Layout), _P, _P, _I,
                           _I, _P], _I),
    "ela_average_down_window": ([ctypes.POINTER(_Layout), ctypes.POINTER(_Layout), _P, _P, ctypes.POINTER(_Layout), _P,
                                 _P, _I, _I, _P], _I),
    "ela_write_plotfile": ([ctypes.c_char_p, _I, _I, _P, _I, _P, _D, _I, _P, _P], _I),
    "ela_server_start": ([_P, _D, _D, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_server_submit": ([_P, _P, ctypes.POINTER(ctypes.c_int64)], _I),
    "ela_server_poll": ([_P, ctypes.c_int64, ctypes.POINTER(_I), _P], _I),
    "ela_server_stop": ([_P, _I], _I),
    "ela_serve": ([_P, _I, ctypes.POINTER(Request), _D, _D, _I, _I, _D, ctypes.POINTER(Result),
                   ctypes.POINTER(ServeSummary)], _I),
    "ela_batch_smoother_times": ([_P, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I)], _I),
    # composite (AMR) solve
    "ela_amr_create_boxes": ([_P, ctypes.POINTER(_Domain), _I, _P, _P, _I, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_amr_create": ([_P, ctypes.POINTER(_Domain), _I, _P, _I, _I, _I, ctypes.POINTER(_P)], _I),
    "ela_amr_destroy": ([_P], _I),
    "ela_amr_bytes": ([_P], _SZ),
    "ela_amr_level_info": ([_P, _I, ctypes.POINTER(_Layout), _P], _I),
    "ela_amr_set_model_field": ([_P, _I, _P], _I),
    "ela_amr_set_bc": ([_P, _P, _P], _I),
    "ela_amr_set_uniform": ([_P, _I], _I),
    "ela_amr_bc_init": ([_P, _I, _P, _I], _I),
    "ela_amr_set_cycle": ([_P, _I, _I, _I], _I),
    "ela_amr_set_level_cycle": ([_P, _I, _I, _I, _I], _I),
    "ela_amr_set_krylov": ([_P, _I], _I),
    "ela_amr_prepare": ([_P], _I),
    "ela_amr_upload": ([_P, _I, _I, _P], _I),
    "ela_amr_download": ([_P, _I, _I, _P], _I),
    "ela_amr_residual": ([_P, ctypes.POINTER(_D)], _I),
    "ela_amr_solve": ([_P, _D, _D, _I, ctypes.POINTER(Stats)], _I),
}


def lib():
    """Load libela.so (raises RuntimeError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libela.so is missing: build it with `python -m anisotropicfracture_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().ela_last_error().decode(errors="replace")
        if rc == 4:
            raise NonConverged(rc, msg)
        raise ElaError(rc, msg)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def ncomp(dim, kind):
    return 2 if kind == ISO else ((21 if dim == 3 else 6) if kind == MM else dim * dim)


class Context:
    """One GPU and an HBM budget (0 = unlimited) that hierarchies are admitted against."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        p = ctypes.c_void_p()
        _check(lib().ela_ctx_create(device, hbm_budget_bytes, ctypes.byref(p)))
        self.ptr = p

    def probe_copy(self, nbytes, nrep=10, variant=0):
        """Average ms of a device copy of nbytes (the HBM-counter calibration probe); variant 0-6
        (ela_probe_copy_ex: more bytes in flight per lane, non-temporal stores, grid sizes)."""
        ms = ctypes.c_double()
        _check(lib().ela_probe_copy_ex(self.ptr, nbytes, nrep, int(variant), ctypes.byref(ms)))
        return ms.value

    def probe_barrier(self, blocks, rounds=2000):
        """(us per device-wide barrier among `blocks` blocks, us per back-to-back empty launch of that grid)
        (ela_probe_barrier)."""
        b, l = ctypes.c_double(), ctypes.c_double()
        _check(lib().ela_probe_barrier(self.ptr, int(blocks), int(rounds), ctypes.byref(b), ctypes.byref(l)))
        return b.value, l.value

    def bytes_in_use(self):
        return lib().ela_ctx_bytes_in_use(self.ptr)

    def close(self):
        if self.ptr:
            lib().ela_ctx_destroy(self.ptr)
            self.ptr = None


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DeviceBuffer:
    """Raw device memory counted against a context's budget (ela_dev_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().ela_dev_alloc(ctx.ptr, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def upload(self, host):
        host = np.ascontiguousarray(host)
        assert host.nbytes <= self.nbytes
        _check(lib().ela_memcpy(self.ptr, _ptr(host), host.nbytes))
        return self

    def download(self, shape, dtype=np.float64):
        out = np.zeros(shape, dtype=dtype)
        _check(lib().ela_memcpy(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().ela_dev_free(self.ctx.ptr, self.ptr)
            self.ptr = None


class DeviceField:
    """A device field on one MG level (owned by the hierarchy)."""

    def __init__(self, op, mglev, prec, nc):
        self.op, self.mglev, self.prec, self.nc = op, mglev, prec, nc
        p = ctypes.c_void_p()
        _check(lib().ela_field_alloc(op.ptr, mglev, prec, nc, ctypes.byref(p)))
        self.ptr = p

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        _check(lib().ela_upload(self.op.ptr, self.mglev, self.prec, self.nc, _ptr(host), self.ptr))
        return self

    def download(self):
        out = np.zeros(self.op.shape(self.mglev, self.nc))
        _check(lib().ela_download(self.op.ptr, self.mglev, self.prec, self.nc, self.ptr, _ptr(out)))
        return out

    def zero(self):
        _check(lib().ela_field_zero(self.op.ptr, self.mglev, self.prec, self.nc, self.ptr))
        return self

    def free(self):
        if self.ptr:
            lib().ela_field_free(self.op.ptr, self.ptr)
            self.ptr = None


class Elastic:
    """Operator::Elastic<T> on one box: ``define`` happens in the constructor.

    n: cells per direction (2 or 3 entries); L: domain lengths (default 1).
    kind: ISO, MM or DIAG (the stiffness storage of the model T).
    mode: SOLVE_FP64, SOLVE_MIXED or ALLOC_BOTH (which MG precisions to allocate).
    """

    def __init__(self, n, kind, L=None, mode=ALLOC_BOTH, max_mg_levels=0, ctx=None, coarse_min=4):
        self.ctx = ctx or default_context()
        n = tuple(int(v) for v in n)
        self.dim = len(n)
        L = L if L is not None else (1.0,) * self.dim
        d = _Domain()
        d.dim = self.dim
        for q in range(3):
            d.n[q] = n[q] if q < self.dim else 0
            d.dx[q] = float(L[q]) / n[q] if q < self.dim else 1.0
        p = ctypes.c_void_p()
        _check(lib().ela_hier_create_ex(self.ctx.ptr, ctypes.byref(d), kind, max_mg_levels, coarse_min, mode,
                                        ctypes.byref(p)))
        self.ptr = p
        self.kind, self.n, self.mode = kind, n, mode
        self.nc = ncomp(self.dim, kind)

    # ------------------------------------------------------------- info
    def levels(self):
        return lib().ela_hier_levels(self.ptr)

    def layout(self, mglev=0):
        lay = _Layout()
        _check(lib().ela_level_info(self.ptr, mglev, ctypes.byref(lay)))
        return lay

    def nodes(self, mglev=0):
        lay = self.layout(mglev)
        return tuple(lay.nodes[q] for q in range(self.dim))

    def shape(self, mglev=0, ncomp=None):
        return (self.dim if ncomp is None else ncomp,) + tuple(reversed(self.nodes(mglev)))

    def bytes(self):
        return lib().ela_hier_bytes(self.ptr)

    def field(self, mglev=0, prec=F64, nc=None):
        return DeviceField(self, mglev, prec, self.dim if nc is None else nc)

    # ------------------------------------------------------------- setup
    def SetModel(self, C):
        """SetModel(FabArray<T>) (Elastic.cpp:84-118): per-node stiffness planes, shape (nc, *nodes)."""
        C = np.ascontiguousarray(C, dtype=np.float64)
        if C.shape != self.shape(0, self.nc):
            raise ValueError("stiffness field has shape %s, expected %s" % (C.shape, self.shape(0, self.nc)))
        _check(lib().ela_set_model_field(self.ptr, _ptr(C)))

    def SetModelVoronoiCubic(self, seeds_q, q, R, C11, C12, C44):
        seeds_q = np.ascontiguousarray(seeds_q, dtype=np.int64)
        R = np.ascontiguousarray(R, dtype=np.float64)
        _check(lib().ela_set_model_voronoi_cubic(self.ptr, seeds_q.shape[0], _ptr(seeds_q), q, _ptr(R), C11, C12, C44))

    def SetModelDegradedIsotropic(self, lambda0, mu0, eta):
        """Degradable::Isotropic::DegradeModulus (Degradable/Isotropic.H:74-104) per node: eta of shape
        (*nodes) or (1, *nodes) degrades lambda and mu by (1 - eta); (2, *nodes) degrades lambda by
        (1 - eta[0]) and mu by (1 - eta[1]).  Raises ElaError (NAN / BADARG) where the reference aborts."""
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        neta = 1 if eta.shape == self.shape(0, 1)[1:] else eta.shape[0]
        if eta.size != neta * int(np.prod(self.shape(0, 1))):
            raise ValueError("eta has shape %s" % (eta.shape,))
        _check(lib().ela_set_model_degraded_iso(self.ptr, lambda0, mu0, neta, _ptr(eta)))

    def _cells_on_device(self, eta_cells):
        cells = np.ascontiguousarray(eta_cells, dtype=np.float64)
        want = tuple(v + 2 for v in reversed(self.n))
        if cells.shape[1:] != want:
            raise ValueError("cell field has shape %s, expected (ncomp,) + %s" % (cells.shape, want))
        buf = DeviceBuffer(self.ctx, cells.nbytes)
        buf.upload(cells)
        return buf, cells.shape[0]

    def DegradeMaterial(self, lambda0, mu0, eta_cells):
        """PolymerDegradation::DegradeMaterial (PolymerDegradation.cpp:793-833) on the device: cell damage of
        shape (number_of_eta, *(n + 2)) with one ghost layer, averaged onto the nodes."""
        buf, neta = self._cells_on_device(eta_cells)
        try:
            _check(lib().ela_degrade_material(self.ptr, lambda0, mu0, neta, buf.ptr, None))
        finally:
            buf.free()

    def CombineCubic(self, ddw, eta_cells, F0=None, want_F0=False):
        """PhaseFieldMicrostructure's fill with Affine::Cubic::Combine (PFM.cpp:529-543, Affine/Cubic.H:70-94):
        ddw (nph, 21) model records, F0 (nph, 3, 3) or None, order parameters (nph, *(n + 2)) in cells.
        Returns the per-node F0 planes (dim*dim, *nodes) when want_F0."""
        ddw = np.ascontiguousarray(ddw, dtype=np.float64)
        F0a = None if F0 is None else np.ascontiguousarray(F0, dtype=np.float64)
        buf, nph = self._cells_on_device(eta_cells)
        fo = self.field(0, F64, self.dim * self.dim) if want_F0 else None
        try:
            _check(lib().ela_combine_cubic(self.ptr, nph, _ptr(ddw), _ptr(F0a) if F0a is not None else None, buf.ptr,
                                           fo.ptr if fo else None, None))
            return fo.download() if fo else None
        finally:
            buf.free()
            if fo:
                fo.free()

    def SetBC(self, types, values=None):
        t = np.ascontiguousarray(types, dtype=np.int32)
        v = np.zeros(t.shape) if values is None else np.ascontiguousarray(values, dtype=np.float64)
        self._bc = (t, v)
        _check(lib().ela_set_bc(self.ptr, _ptr(t), _ptr(v)))

    def SetUniform(self, uniform):
        _check(lib().ela_set_uniform(self.ptr, int(bool(uniform))))

    def SetCycle(self, cycle=CYCLE_F, nu1=2, nu2=2):
        _check(lib().ela_set_cycle(self.ptr, cycle, nu1, nu2))

    def SetStiffnessMode(self, mode):
        """0: per-node stiffness planes; 1: per-node grain id + per-grain records (Voronoi models)."""
        _check(lib().ela_set_stiffness_mode(self.ptr, int(mode)))

    def set_option(self, option, value):
        """ela_set_option: OPT_MARCH (LDS z-marching kernels for grain-form stiffness, default on),
        OPT_KZ (planes per marching tile, default 16), OPT_XCD (tile order) or OPT_MASK (plane levels
        skip the derivative of C where it is exactly zero, constant-C nodes launched first: 0 never,
        1 where it pays, 2 always with node lists, 3 always with in-block regrouping) or OPT_GRAPH (the cycles'
        launches on levels >= 1 replayed from captured HIP graphs, default off) or OPT_FOLD (the last sweep of a
        Krylov step writes the FP64 vector itself, default on; the same bits either way) or OPT_ROWS (rows per
        thread of level 0's FP32 marching tiles, 1 or 2; the same bits either way) or OPT_PREFETCH (planes a marching
        tile loads ahead, 1 or 2; the same bits either way) or OPT_GJ (the dense coarsest inverse by Gauss-Jordan in
        one launch a column, default 1, or two; the same bits either way) or OPT_GRAIN_DIAG (marching tiles of a
        grain-form level take the interior Jacobi diagonal from the node's grain, default 1, or from the stored field;
        the same bits either way)."""
        _check(lib().ela_set_option(self.ptr, option, int(value)))

    def autotune(self, nrep=10):
        """Pick the marching tile height / order / rows per thread / planes loaded ahead by timing level-0
        sweeps (ela_autotune); returns {(kz, xcd, rows, prefetch): ms} (rows 2 only for FP32 sweeps)."""
        ms = np.zeros(24)
        _check(lib().ela_autotune(self.ptr, int(nrep), _ptr(ms)))
        idx = lambda pf, ry, a, x: 12 * (pf - 1) + 6 * (ry - 1) + 2 * a + x
        return {(kz, x, ry, pf): ms[idx(pf, ry, a, x)] for pf in (1, 2) for ry in (1, 2)
                for a, kz in enumerate((8, 16, 32)) for x in (0, 1) if ms[idx(pf, ry, a, x)] > 0}

    def SetKrylov(self, restart):
        """Flexible GMRES(restart) around the MG cycle (0: plain MG iteration)."""
        _check(lib().ela_set_krylov(self.ptr, int(restart)))

    def prepareForSolve(self):
        _check(lib().ela_prepare(self.ptr))

    def bc_init(self, rhs=None, homogeneous=False):
        rhs = np.zeros(self.shape(0)) if rhs is None else np.ascontiguousarray(rhs, dtype=np.float64).copy()
        _check(lib().ela_bc_init(self.ptr, _ptr(rhs), int(homogeneous)))
        return rhs

    # ------------------------------------------------------------- operator (host arrays in, host arrays out)
    def Fapply(self, u, mglev=0, prec=F64):
        fu = self.field(mglev, prec).upload(u)
        ff = self.field(mglev, prec)
        _check(lib().ela_fapply(self.ptr, mglev, prec, fu.ptr, ff.ptr, None))
        out = ff.download()
        fu.free()
        ff.free()
        return out

    def newton_rhs(self, u, b, F0=None):
        """Newton<T>::prepareForSolve's right-hand side (Newton.H:30-140): b - div(DDW : (grad u - F0)) inside,
        b - BC row(u, grad u, dw) on the boundary.  F0: (dim*dim, *nodes) per-node eigenstrain or None."""
        fu = self.field(0, F64).upload(u)
        fb = self.field(0, F64).upload(b)
        fr = self.field(0, F64)
        ff = self.field(0, F64, self.dim * self.dim).upload(F0) if F0 is not None else None
        _check(lib().ela_newton_rhs(self.ptr, fu.ptr, fb.ptr, ff.ptr if ff else None, fr.ptr, None))
        out = fr.download()
        for f in (fu, fb, fr, ff):
            if f:
                f.free()
        return out

    def Diagonal(self, mglev=0, prec=F64):
        fd = self.field(mglev, prec)
        _check(lib().ela_diagonal(self.ptr, mglev, prec, fd.ptr, None))
        out = fd.download()
        fd.free()
        return out

    def _post(self, what, u, voigt):
        nc = 1 if what == 2 else ((6 if self.dim == 3 else 3) if voigt else self.dim * self.dim)
        fu = self.field(0, F64, self.dim).upload(u)
        fo = self.field(0, F64, nc)
        _check(lib().ela_postprocess(self.ptr, what, int(bool(voigt)), fu.ptr, fo.ptr, None))
        out = fo.download()
        fu.free()
        fo.free()
        return out

    def W_DW(self, u, F0=None, wfac=0.5):
        """Newton<T>::W and DW (Newton.H:233-333): energy density (*nodes) and dw = DDW : (grad u - F0)
        (dim*dim, *nodes).  wfac 0.5 for the Linear models and Affine::Cubic, 1 for Affine::Isotropic."""
        fu = self.field(0, F64).upload(u)
        fw = self.field(0, F64, 1)
        fd = self.field(0, F64, self.dim * self.dim)
        ff = self.field(0, F64, self.dim * self.dim).upload(F0) if F0 is not None else None
        _check(lib().ela_newton_w_dw(self.ptr, fu.ptr, ff.ptr if ff else None, wfac, fw.ptr, fd.ptr, None))
        w, dw = fw.download()[0], fd.download()
        for f in (fu, fw, fd, ff):
            if f:
                f.free()
        return w, dw

    def Strain(self, u, voigt=False):
        """Elastic<T>::Strain (Elastic.cpp:399-447): sym(grad u) per node."""
        return self._post(0, u, voigt)

    def Stress(self, u, voigt=False):
        """Elastic<T>::Stress (Elastic.cpp:451-499): C : grad u per node."""
        return self._post(1, u, voigt)

    def Energy(self, u):
        """Elastic<T>::Energy (Elastic.cpp:502-549): W(grad u) + sigma:eps/2 per node (the reference's sum)."""
        return self._post(2, u, False)[0]

    def normalize(self, x, mglev=0, prec=F64):
        """Operator<Node>::normalize (Operator.cpp:164-195) as the reference computes it: component 0
        of x divided by diagonal component 0 once per component; other components unchanged."""
        f = self.field(mglev, prec).upload(x)
        _check(lib().ela_normalize(self.ptr, mglev, prec, f.ptr, None))
        out = f.download()
        f.free()
        return out

    def Fsmooth(self, x, b, mglev=0, prec=F64, nsweeps=2, omega=2.0 / 3.0):
        fx = self.field(mglev, prec).upload(x)
        fb = self.field(mglev, prec).upload(b)
        _check(lib().ela_fsmooth(self.ptr, mglev, prec, fx.ptr, fb.ptr, nsweeps, omega, None))
        out = fx.download()
        fx.free()
        fb.free()
        return out

    def restriction(self, fine, cmglev, prec=F64):
        ff = self.field(cmglev - 1, prec).upload(fine)
        fc = self.field(cmglev, prec)
        _check(lib().ela_restriction(self.ptr, cmglev, prec, ff.ptr, fc.ptr, None))
        out = fc.download()
        ff.free()
        fc.free()
        return out

    def interpolation(self, fine, crse, fmglev, prec=F64):
        ff = self.field(fmglev, prec).upload(fine)
        fc = self.field(fmglev + 1, prec).upload(crse)
        _check(lib().ela_interpolation(self.ptr, fmglev, prec, ff.ptr, fc.ptr, None))
        out = ff.download()
        ff.free()
        fc.free()
        return out

    def residual(self, x, b, mglev=0, prec=F64):
        fx = self.field(mglev, prec).upload(x)
        fb = self.field(mglev, prec).upload(b)
        fr = self.field(mglev, prec)
        nrm = ctypes.c_double()
        _check(lib().ela_residual(self.ptr, mglev, prec, fx.ptr, fb.ptr, fr.ptr, ctypes.byref(nrm), None))
        out = fr.download()
        for f in (fx, fb, fr):
            f.free()
        return out, nrm.value

    def solve_device(self, b_field, u_field, tol_rel, tol_abs=0.0, mode=SOLVE_FP64, max_iter=200):
        """Linear::solve on device-resident level-0 FP64 fields (u holds the initial guess)."""
        st = Stats()
        rc = lib().ela_solve_device(self.ptr, b_field.ptr, u_field.ptr, tol_rel, tol_abs, mode, max_iter,
                                    ctypes.byref(st))
        return rc, st.as_dict()

    def time_fapply(self, prec=F64, nrep=20):
        ms = ctypes.c_double()
        _check(lib().ela_time_fapply(self.ptr, prec, nrep, ctypes.byref(ms)))
        return ms.value

    def set_timing(self, on=True):
        _check(lib().ela_set_timing(self.ptr, int(bool(on))))

    def kernel_times(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        na, nb = ctypes.c_int(), ctypes.c_int()
        _check(lib().ela_kernel_times(self.ptr, ctypes.byref(a), ctypes.byref(na), ctypes.byref(b), ctypes.byref(nb)))
        return {"resid64_ms": a.value, "resid64_n": na.value, "smooth0_ms": b.value, "smooth0_n": nb.value}

    def close(self):
        if getattr(self, "ptr", None):
            lib().ela_hier_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Linear:
    """Solver::Nonlocal::Linear (Linear.H:19-143): holds a reference to the operator."""

    def __init__(self, op):
        self.op = op
        self.tol_rel, self.tol_abs = 1e-8, 1e-8  # Linear.H:118
        self.max_iter = 200
        self.last = None

    def solve(self, sol, rhs, tol_rel=None, tol_abs=None, mode=SOLVE_FP64, max_iter=None, raise_on_fail=True):
        """Solve A sol = rhs from the initial guess in sol; returns (sol, stats dict)."""
        tol_rel = self.tol_rel if tol_rel is None else tol_rel
        tol_abs = self.tol_abs if tol_abs is None else tol_abs
        max_iter = self.max_iter if max_iter is None else max_iter
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        u = np.ascontiguousarray(sol, dtype=np.float64).copy()
        st = Stats()
        rc = lib().ela_solve(self.op.ptr, _ptr(b), _ptr(u), tol_rel, tol_abs, mode, max_iter, ctypes.byref(st))
        self.last = st.as_dict()
        if rc != 0 and (raise_on_fail or rc not in (4,)):
            _check(rc)
        return u, self.last


class Newton(Linear):
    """Solver::Nonlocal::Newton<T> (Newton.H:15-349) for the models whose kinematic variable is grad u:
    each of nriters steps builds rhs = b - div(DDW : (grad u - F0)) (prepareForSolve), solves A du = rhs
    from du = 0 and adds du to u.  F0 (dim*dim planes per node, or None) makes the models affine
    (Affine::Isotropic / Affine::Cubic, Affine/*.H)."""

    def __init__(self, op, nriters=1):
        super().__init__(op)
        self.nriters = nriters  # Newton.H:337

    def solve(self, u, b, F0=None, tol_rel=None, tol_abs=None, mode=SOLVE_FP64, max_iter=None, raise_on_fail=True):
        u = np.array(u, dtype=np.float64, copy=True)
        st = None
        for _ in range(self.nriters):
            rhs = self.op.newton_rhs(u, b, F0)
            du, st = Linear.solve(self, np.zeros_like(u), rhs, tol_rel, tol_abs, mode, max_iter, raise_on_fail)
            u += du
        return u, st


# ---------------------------------------------------------------- AMR level transfers
def _pair_args(crse, fine, fine_origin):
    if crse.prec != fine.prec or crse.nc != fine.nc:
        raise ValueError("coarse and fine fields differ in precision or components")
    lc, lf = crse.op.layout(crse.mglev), fine.op.layout(fine.mglev)
    o = np.zeros(3, dtype=np.int32)
    o[:len(fine_origin)] = fine_origin
    return lc, lf, o


def reflux(crse_res, fine_res, fine_origin):
    """Operator<Node>::reflux (Operator.cpp:728-839) between two device fields: the coarse residual
    (a field of the coarse level's hierarchy) takes the full-weighting restriction of the fine
    residual (a field on the fine box's hierarchy) on the coarsened fine box.  fine_origin: fine-level
    node index (x, y[, z]) of the fine box's first node."""
    lc, lf, o = _pair_args(crse_res, fine_res, fine_origin)
    _check(lib().ela_reflux(ctypes.byref(lc), crse_res.ptr, ctypes.byref(lf), fine_res.ptr, _ptr(o),
                            crse_res.nc, crse_res.prec, None))


def average_down(crse_sol, fine_sol, fine_origin):
    """averageDownSolutionRHS (Operator.cpp:626-641): coarse nodes of the coarsened fine box take the
    coincident fine solution value."""
    lc, lf, o = _pair_args(crse_sol, fine_sol, fine_origin)
    _check(lib().ela_average_down(ctypes.byref(lc), crse_sol.ptr, ctypes.byref(lf), fine_sol.ptr, _ptr(o),
                                  crse_sol.nc, crse_sol.prec, None))


# ---------------------------------------------------------------- composite (AMR) solve
class ElasticAMR:
    """Operator::Elastic<T> defined on several AMR levels (Elastic.cpp:30-55) with the multi-level
    Solver::Nonlocal::Linear::solve (Linear.H:102-112).

    n0: level-0 cells per direction; boxes: one entry per level 1.., either a box (lo, hi) or a list of
    boxes [(lo, hi), ...] whose union is the level (node indices of the level, ratio 2, even corners;
    e.g. a region chopped into boxes of at most max_grid_size cells, or an L shape).  Every per-level
    host array covers the bounding box of the level's union grown by 2 nodes and clipped to the domain
    (``shape(lev)``; ``origin(lev)`` is the global node index of its first node); nodes outside the
    union and on its coarse/fine boundary hold the interpolated coarser solution.
    """

    U, B, R = 0, 1, 2

    def __init__(self, n0, boxes, kind, L=None, mode=SOLVE_MIXED, coarse_min=4, ctx=None):
        self.ctx = ctx or default_context()
        n0 = tuple(int(v) for v in n0)
        self.dim = len(n0)
        L = L if L is not None else (1.0,) * self.dim
        d = _Domain()
        d.dim = self.dim
        for q in range(3):
            d.n[q] = n0[q] if q < self.dim else 0
            d.dx[q] = float(L[q]) / n0[q] if q < self.dim else 1.0
        self.nlev = len(boxes) + 1
        lists = [[b] if np.ndim(b[0]) == 1 else list(b) for b in boxes]
        counts = np.array([len(b) for b in lists] or [0], dtype=np.int32)
        flat = [bb for b in lists for bb in b]
        bx = np.zeros(6 * max(1, len(flat)), dtype=np.int32)
        for q, (lo, hi) in enumerate(flat):
            bx[6 * q:6 * q + self.dim] = lo
            bx[6 * q + 3:6 * q + 3 + self.dim] = hi
        self.boxes = lists
        self.n0 = n0
        p = ctypes.c_void_p()
        _check(lib().ela_amr_create_boxes(self.ctx.ptr, ctypes.byref(d), self.nlev, _ptr(counts), _ptr(bx), kind,
                                          coarse_min, mode, ctypes.byref(p)))
        self.ptr = p
        self.kind, self.mode = kind, mode
        self.nc = ncomp(self.dim, kind)

    def WritePlotFile(self, path, fields, time=0.0, step=0):
        """The levels' node fields as a plot file (Integrator::WritePlotFile, Integrator.cpp:582-687):
        fields maps a name to per-level arrays (ncomp, *shape(lev)); a field of several components is
        written as name001, name002, ... as the reference names them."""
        names, per = [], [[] for _ in range(self.nlev)]
        for name, arrs in fields.items():
            nc = arrs[0].shape[0]
            names += [name] if nc == 1 else ["%s%03d" % (name, j + 1) for j in range(nc)]
            for l in range(self.nlev):
                per[l].append(np.asarray(arrs[l]).reshape((nc,) + tuple(arrs[l].shape[1:])))
        levels = []
        for l in range(self.nlev):
            data = np.concatenate(per[l], axis=0)
            dom = [n << l for n in self.n0]
            boxes = [((0,) * self.dim, tuple(dom))] if l == 0 else self.boxes[l - 1]
            levels.append({"domain": dom, "origin": self.origin(l), "data": data, "boxes": boxes})
        write_plotfile(path, levels, names, time, step)

    def _info(self, lev):
        lay = _Layout()
        o = np.zeros(3, dtype=np.int32)
        _check(lib().ela_amr_level_info(self.ptr, lev, ctypes.byref(lay), _ptr(o)))
        return lay, o

    def shape(self, lev, ncomp=None):
        lay, _ = self._info(lev)
        nodes = tuple(lay.nodes[q] for q in range(self.dim))
        return ((self.dim if ncomp is None else ncomp),) + tuple(reversed(nodes))

    def origin(self, lev):
        return tuple(int(v) for v in self._info(lev)[1][:self.dim])

    def bytes(self):
        return lib().ela_amr_bytes(self.ptr)

    def SetModel(self, lev, C):
        C = np.ascontiguousarray(C, dtype=np.float64)
        if C.shape != self.shape(lev, self.nc):
            raise ValueError("model shape %s, expected %s" % (C.shape, self.shape(lev, self.nc)))
        _check(lib().ela_amr_set_model_field(self.ptr, lev, _ptr(C)))

    def SetBC(self, types, values=None):
        t = np.ascontiguousarray(types, dtype=np.int32)
        v = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().ela_amr_set_bc(self.ptr, _ptr(t), None if v is None else _ptr(v)))

    def SetUniform(self, uniform):
        _check(lib().ela_amr_set_uniform(self.ptr, int(bool(uniform))))

    def bc_init(self, lev, rhs=None, homogeneous=False):
        """BC::Operator::Elastic::Init on level lev's array (domain-boundary nodes take their values)."""
        out = np.zeros(self.shape(lev)) if rhs is None else np.array(rhs, dtype=np.float64, copy=True)
        _check(lib().ela_amr_bc_init(self.ptr, lev, _ptr(out), int(bool(homogeneous))))
        return out

    def coords(self, lev):
        """Nodal coordinates of level lev's array, one array per direction."""
        lay, o = self._info(lev)
        axes = [(o[d] + np.arange(lay.nodes[d])) * lay.dx[d] for d in range(self.dim)]
        return list(reversed(np.meshgrid(*reversed(axes), indexing="ij")))

    def SetCycle(self, cycle=CYCLE_F, nu1=2, nu2=2, lev=None):
        if lev is None:
            _check(lib().ela_amr_set_cycle(self.ptr, cycle, nu1, nu2))
        else:
            _check(lib().ela_amr_set_level_cycle(self.ptr, lev, cycle, nu1, nu2))

    def SetKrylov(self, restart):
        _check(lib().ela_amr_set_krylov(self.ptr, restart))

    def prepareForSolve(self):
        _check(lib().ela_amr_prepare(self.ptr))

    def upload(self, field, lev, host):
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.shape != self.shape(lev):
            raise ValueError("field shape %s, expected %s" % (a.shape, self.shape(lev)))
        _check(lib().ela_amr_upload(self.ptr, lev, field, _ptr(a)))

    def download(self, field, lev):
        out = np.zeros(self.shape(lev))
        _check(lib().ela_amr_download(self.ptr, lev, field, _ptr(out)))
        return out

    def residual(self, u, b):
        """Composite residual (solutionResidual + reflux) of per-level u, b; (r per level, |r|inf)."""
        for l in range(self.nlev):
            self.upload(self.U, l, u[l])
            self.upload(self.B, l, b[l])
        nrm = ctypes.c_double()
        _check(lib().ela_amr_residual(self.ptr, ctypes.byref(nrm)))
        return [self.download(self.R, l) for l in range(self.nlev)], nrm.value

    def solve_resident(self, tol_rel=1e-10, tol_abs=0.0, max_iter=200, raise_on_fail=True):
        st = Stats()
        rc = lib().ela_amr_solve(self.ptr, tol_rel, tol_abs, max_iter, ctypes.byref(st))
        if rc != 0 and (raise_on_fail or rc != 4):
            _check(rc)
        return st.as_dict()

    def solve(self, sol, rhs, tol_rel=1e-10, tol_abs=0.0, max_iter=200, raise_on_fail=True):
        """Solve from the initial guess sol (per-level arrays); returns (sol per level, stats dict)."""
        for l in range(self.nlev):
            self.upload(self.U, l, sol[l])
            self.upload(self.B, l, rhs[l])
        st = self.solve_resident(tol_rel, tol_abs, max_iter, raise_on_fail)
        return [self.download(self.U, l) for l in range(self.nlev)], st

    def close(self):
        if getattr(self, "ptr", None):
            lib().ela_amr_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- batched independent solves
def _ptr_array(objs):
    arr = (ctypes.c_void_p * len(objs))(*[o.ptr for o in objs])
    return arr


class RVE:
    """One member of a :class:`Batch`: a cube of n cells on [0,1]^3 (cubic
    Voronoi polycrystal), its hierarchy resident on the device.

    Dense host fields are ``(3, n+1, n+1, n+1)`` float64 arrays (x fastest)."""

    U, B, R, FACTOR = 0, 1, 2, 3

    def __init__(self, batch, n, max_grains=64, with_factor=False):
        self.batch = batch
        p = ctypes.c_void_p()
        _check(lib().ela_rve_create(batch.ptr, int(n), int(max_grains), int(bool(with_factor)), ctypes.byref(p)))
        self.ptr = p
        self.n = int(n)
        batch._members.append(self)

    @property
    def shape(self):
        return (3, self.n + 1, self.n + 1, self.n + 1)

    def levels(self):
        return lib().ela_rve_levels(self.ptr)

    def bytes(self):
        return lib().ela_rve_bytes_in_use(self.ptr)

    def SetBC(self, types):
        t = np.ascontiguousarray(types, dtype=np.int32).reshape(26, 3)
        _check(lib().ela_rve_set_bc_types(self.ptr, _ptr(t)))

    def SetUniform(self, uniform):
        _check(lib().ela_rve_set_uniform(self.ptr, int(bool(uniform))))

    def SetModelVoronoiCubic(self, seeds_q, q, R, C11, C12, C44):
        sd = np.ascontiguousarray(seeds_q, dtype=np.int64)
        Rm = np.ascontiguousarray(R, dtype=np.float64)
        _check(lib().ela_rve_set_voronoi_cubic(self.ptr, sd.shape[0], _ptr(sd), int(q), _ptr(Rm), C11, C12, C44))

    def SetCrack(self, p, nrm, tv, width, a, b, c):
        pa, na, ta = (np.ascontiguousarray(v, dtype=np.float64) for v in (p, nrm, tv))
        _check(lib().ela_rve_set_crack(self.ptr, _ptr(pa), _ptr(na), _ptr(ta), width, a, b, c))

    def SetRhsAffine(self, G):
        g = np.ascontiguousarray(G, dtype=np.float64).reshape(9)
        _check(lib().ela_rve_set_rhs_affine(self.ptr, _ptr(g)))

    def upload(self, field, host):
        h = np.ascontiguousarray(host, dtype=np.float64).reshape(self.shape)
        _check(lib().ela_rve_upload(self.ptr, field, _ptr(h)))

    def download(self, field):
        shape = self.shape[1:] if field == self.FACTOR else self.shape
        out = np.empty(shape)
        _check(lib().ela_rve_download(self.ptr, field, _ptr(out)))
        return out

    def grain_ids(self):
        out = np.empty(self.shape[1:], dtype=np.uint16)
        _check(lib().ela_rve_grain_ids(self.ptr, _ptr(out)))
        return out

    def status(self):
        st, code = ctypes.c_int(), ctypes.c_int()
        s = Stats()
        _check(lib().ela_rve_status(self.ptr, ctypes.byref(st), ctypes.byref(code), ctypes.byref(s)))
        d = s.as_dict()
        d.update({"state": st.value, "status": code.value})
        return d

    def close(self):
        if getattr(self, "ptr", None):
            lib().ela_rve_destroy(self.ptr)
            self.ptr = None
            if self in self.batch._members:
                self.batch._members.remove(self)


def write_plotfile(path, levels, names, time=0.0, step=0, prob_lo=None, prob_hi=None):
    """ela_write_plotfile: a nodal plot file in AMReX's layout (Integrator::WritePlotFile,
    Integrator.cpp:582-687).  levels: dicts with domain (cells per direction), origin (node index of the
    array's first node), data (ncomp, [nz,] ny, nx) float64 and boxes [(lo, hi), ...] of nodes."""
    dim = len(levels[0]["domain"])
    arr = (PlotLevel * len(levels))()
    hold = []
    for l, lv in enumerate(levels):
        data = np.ascontiguousarray(lv["data"], dtype=np.float64)
        if data.shape[0] != len(names):
            raise ValueError("level %d has %d components for %d names" % (l, data.shape[0], len(names)))
        bx = np.zeros(6 * len(lv["boxes"]), dtype=np.int32)
        for q, (lo, hi) in enumerate(lv["boxes"]):
            bx[6 * q:6 * q + dim] = lo
            bx[6 * q + 3:6 * q + 3 + dim] = hi
        hold += [data, bx]
        a = arr[l]
        nodes = list(reversed(data.shape[1:])) + [1] * (3 - dim)
        for d in range(3):
            a.domain[d] = int(lv["domain"][d]) if d < dim else 0
            a.origin[d] = int(lv.get("origin", (0,) * dim)[d]) if d < dim else 0
            a.nodes[d] = int(nodes[d])
        a.data, a.nboxes, a.boxes = data.ctypes.data, len(lv["boxes"]), bx.ctypes.data
    cn = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    lo = np.ascontiguousarray(prob_lo if prob_lo is not None else np.zeros(dim), dtype=np.float64)
    hi = np.ascontiguousarray(prob_hi if prob_hi is not None else np.ones(dim), dtype=np.float64)
    _check(lib().ela_write_plotfile(str(path).encode(), dim, len(levels), arr, len(names), cn, float(time),
                                    int(step), _ptr(lo), _ptr(hi)))


def _fill_request(r, q):
    """An ela_request from a request dict (arrival, n, seeds, R, G; optional crack, C11/C12/C44,
    max_cycles); returns the arrays it points into (keep them alive while the request is in use)."""
    from . import configs as K
    sd = np.ascontiguousarray(q["seeds"], dtype=np.int64)
    Rm = np.ascontiguousarray(q["R"], dtype=np.float64).reshape(-1, 9)
    r.arrival, r.n, r.nseeds, r.q = float(q.get("arrival", 0.0)), int(q["n"]), sd.shape[0], K.VORONOI_Q
    r.seeds, r.R = sd.ctypes.data, Rm.ctypes.data
    r.C11, r.C12, r.C44 = q.get("C11", K.C11), q.get("C12", K.C12), q.get("C44", K.C44)
    cr = q.get("crack")
    r.crack = 1 if cr is not None else 0
    if cr is not None:
        p, nrm, tv, width, g = cr
        r.crack_p[:], r.crack_nrm[:], r.crack_tv[:] = list(p), list(nrm), list(tv)
        r.crack_width = float(width)
        r.crack_g[:] = list(g)
    r.G[:] = list(np.asarray(q["G"], dtype=np.float64).reshape(9))
    r.max_cycles = int(q.get("max_cycles", 0))
    return [sd, Rm]


class Server:
    """Online serving on a batch (ela_server_*): a host thread admits submitted requests into the running
    continuous-batching loop.  submit() returns a ticket; poll() returns None until the result is ready."""

    def __init__(self, batch, tol_rel=1e-10, tol_abs=0.0, max_cycles=100, max_active=256):
        p = ctypes.c_void_p()
        _check(lib().ela_server_start(batch.ptr, tol_rel, tol_abs, max_cycles, max_active, ctypes.byref(p)))
        self.ptr, self.batch, self._hold = p, batch, {}

    def submit(self, q, u_out=None):
        r = Request()
        arrays = _fill_request(r, q)
        if u_out is not None:
            r.u_out = u_out.ctypes.data
            arrays.append(u_out)
        t = ctypes.c_int64()
        _check(lib().ela_server_submit(self.ptr, ctypes.byref(r), ctypes.byref(t)))
        self._hold[t.value] = arrays
        return t.value

    def poll(self, ticket):
        done, res = ctypes.c_int(), Result()
        _check(lib().ela_server_poll(self.ptr, ctypes.c_int64(ticket), ctypes.byref(done), ctypes.byref(res)))
        if not done.value:
            return None
        self._hold.pop(ticket, None)
        return {k: getattr(res, k) for k, _ in Result._fields_}

    def stop(self, drain=True):
        if self.ptr:
            p, self.ptr = self.ptr, None
            _check(lib().ela_server_stop(p, int(bool(drain))))


class Batch:
    """Batched independent solves (mode SOLVE_MIXED or SOLVE_FP64)."""

    def __init__(self, mode=SOLVE_MIXED, ctx=None):
        self.ctx = ctx or default_context()
        p = ctypes.c_void_p()
        _check(lib().ela_batch_create(self.ctx.ptr, mode, ctypes.byref(p)))
        self.ptr = p
        self.mode = mode
        self._members = []

    @staticmethod
    def member_bytes(n, mode=SOLVE_MIXED, max_grains=64, with_factor=False):
        return lib().ela_rve_bytes(int(n), mode, int(max_grains), int(bool(with_factor)))

    def rve(self, n, max_grains=64, with_factor=False):
        return RVE(self, n, max_grains, with_factor)

    def prepare(self, rves):
        _check(lib().ela_batch_prepare(self.ptr, len(rves), _ptr_array(rves)))

    def fapply(self, rves, prec=F64):
        _check(lib().ela_batch_fapply(self.ptr, len(rves), _ptr_array(rves), prec))

    def submit(self, rve, tol_rel=1e-10, tol_abs=0.0, max_cycles=100):
        _check(lib().ela_batch_submit(self.ptr, rve.ptr, tol_rel, tol_abs, max_cycles))

    def step(self):
        n = ctypes.c_int()
        _check(lib().ela_batch_step(self.ptr, ctypes.byref(n)))
        return n.value

    def solve(self, rves, tol_rel=1e-10, tol_abs=0.0, max_cycles=100, raise_on_fail=True):
        rc = lib().ela_batch_solve(self.ptr, len(rves), _ptr_array(rves), tol_rel, tol_abs, max_cycles)
        if rc != 0 and (raise_on_fail or rc != 4):
            _check(rc)
        return [r.status() for r in rves]

    def set_timing(self, on=True):
        _check(lib().ela_batch_set_timing(self.ptr, int(bool(on))))

    MARCH, KZ = 0, 1

    def set_option(self, option, value):
        """ela_batch_set_option: OPT_MARCH (LDS z-marching kernels on the grain-form levels, default on),
        OPT_KZ (planes per marching tile, default 16), OPT_XCD (tile order), OPT_GRAPH (each step's F-cycle
        replayed from a captured HIP graph while the active set is unchanged, default off) or OPT_GJ (the members'
        coarsest Gauss-Jordan in one launch a column, default 1, or two; the same bits) or OPT_GRAIN_DIAG (the
        marching tiles of members without factors take the interior Jacobi diagonal per grain, default 1, or read
        the stored field; the same bits) or OPT_PROLONG (prolongation with one thread per coarse node, default 1, or
        per fine node; the same bits)."""
        _check(lib().ela_batch_set_option(self.ptr, option, int(value)))

    def graph_stats(self):
        """(captures, replays) of the F-cycle graph (OPT_GRAPH)."""
        c, r = ctypes.c_int(), ctypes.c_int()
        _check(lib().ela_batch_graph_stats(self.ptr, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def autotune(self, rves, nrep=5):
        """ela_batch_autotune over the finest levels of `rves`; returns {(kz, xcd): ms}."""
        ms = np.zeros(6)
        _check(lib().ela_batch_autotune(self.ptr, len(rves), _ptr_array(rves), int(nrep), _ptr(ms)))
        return {(kz, x): ms[2 * a + x] for a, kz in enumerate((8, 16, 32)) for x in (0, 1)}

    def serve(self, requests, tol_rel=1e-10, tol_abs=0.0, max_cycles=100, max_active=256, max_seconds=0.0,
              keep=()):
        """Run the request loop (ela_serve) over `requests` (dicts: arrival, n, seeds, R, G and optional
        crack = (p, nrm, tv, width, (g2, g1, g0))), sorted by arrival.  Solutions of the request indices in
        `keep` are returned.  Returns (results: list of dicts, solutions: {index: array})."""
        reqs = (Request * len(requests))()
        hold, sols = [], {}
        for i, q in enumerate(requests):
            hold += _fill_request(reqs[i], q)
            if i in keep:
                sols[i] = np.zeros((3,) + (int(q["n"]) + 1,) * 3)
                reqs[i].u_out = sols[i].ctypes.data
        res = (Result * len(requests))()
        summ = ServeSummary()
        _check(lib().ela_serve(self.ptr, len(requests), reqs, tol_rel, tol_abs, max_cycles, max_active, max_seconds,
                               res, ctypes.byref(summ)))
        out = [{k: getattr(x, k) for k, _ in Result._fields_} for x in res]
        self.last_serve = {k: getattr(summ, k) for k, _ in ServeSummary._fields_}
        return out, sols

    def smoother_times(self):
        ms, nn, nl = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_int()
        _check(lib().ela_batch_smoother_times(self.ptr, ctypes.byref(ms), ctypes.byref(nn), ctypes.byref(nl)))
        return {"ms": ms.value, "nodes": nn.value, "launches": nl.value}

    def bytes(self):
        return lib().ela_batch_bytes_in_use(self.ptr)

    def close(self):
        """Destroy the batch and every member still in it."""
        if getattr(self, "ptr", None):
            lib().ela_batch_destroy(self.ptr)
            self.ptr = None
            for r in self._members:
                r.ptr = None
            self._members = []
