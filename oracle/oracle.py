"""ctypes wrapper of the CPU FP64 oracle (oracle/elastic_oracle.cpp).

Test infrastructure only: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's ``cpu_baseline`` leg, always as the checker.  The product
package never imports it.

Field layout: component-major, x fastest -- ``v[c, k, j, i]`` (3D) or
``v[c, j, i]`` (2D), i.e. a numpy array of shape ``(ncomp, *reversed(nodes))``.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "elastic_oracle.cpp")
LIB = os.path.join(HERE, "_build", "libelastic_oracle.so")

ISO, MM, DIAG = 0, 1, 2
DISP, TRAC, PERIODIC, NEUMANN = 0, 1, 2, 3

# Face numbering of the reference's BC::Operator::Elastic::Face enum
# (src/BC/Operator/Elastic.H:24-42).
FACES3 = ["XLO", "YLO", "ZLO", "XHI", "YHI", "ZHI",
          "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI",
          "ZLO_XLO", "ZLO_XHI", "ZHI_XLO", "ZHI_XHI",
          "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI",
          "XLO_YLO_ZLO", "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI",
          "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI"]
FACES2 = ["XLO", "YLO", "XHI", "YHI", "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI"]


def build(force=False):
    """Compile the oracle with g++ (OpenMP) into oracle/_build/."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17",
           "-o", LIB + ".tmp", SRC]
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


class _Grid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n", ctypes.c_int * 3), ("dx", ctypes.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER
        d, i = ctypes.c_double, ctypes.c_int
        vp = ctypes.c_void_p
        L.or_ncomp.argtypes = [i, i]
        L.or_mm_slot.argtypes = [i, i, i, i, i]
        L.or_classify.argtypes = [P(_Grid), i, i, i, vp]
        L.or_fapply.argtypes = [P(_Grid), i, vp, vp, i, vp, vp]
        L.or_diagonal.argtypes = [P(_Grid), i, vp, vp, vp]
        L.or_restrict.argtypes = [P(_Grid), i, vp, vp]
        L.or_interp_add.argtypes = [P(_Grid), i, vp, vp]
        L.or_fsmooth.argtypes = [P(_Grid), i, vp, vp, vp, vp, vp]
        L.or_bc_init.argtypes = [P(_Grid), vp, vp, i, vp]
        L.or_cubic_define_R.argtypes = [i, d, d, d, vp, vp]
        L.or_cubic_rotation.argtypes = [d, d, d, vp]
        L.or_assemble.argtypes = [P(_Grid), i, vp, vp, vp]
        L.or_mg_solve.argtypes = [P(_Grid), i, vp, vp, vp, vp, d, d, i, i, i, i, i, vp, vp, i]
        L.or_mg_solve.restype = i
        L.or_reflux.argtypes = [P(_Grid), i, vp, vp, vp, vp]
        L.or_normalize.argtypes = [P(_Grid), vp, vp]
        L.or_post.argtypes = [P(_Grid), i, vp, vp, i, i, vp]
        L.or_newton_rhs.argtypes = [P(_Grid), i, vp, vp, vp, vp, vp, vp]
        L.or_newton_w_dw.argtypes = [P(_Grid), i, vp, vp, vp, d, vp, vp]
        L.or_set_threads.argtypes = [i]
        L.or_get_threads.restype = i
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Grid:
    """A single-box nodal grid of n cells per direction on [0, L]^dim."""

    def __init__(self, n, L=None):
        n = tuple(int(v) for v in n)
        self.dim = len(n)
        self.n = n
        L = L if L is not None else (1.0,) * self.dim
        self.dx = tuple(float(L[d]) / n[d] for d in range(self.dim))
        self.nodes = tuple(v + 1 for v in n)
        self.nnodes = int(np.prod(self.nodes))

    def c(self):
        g = _Grid()
        g.dim = self.dim
        for d in range(3):
            g.n[d] = self.n[d] if d < self.dim else 0
            g.dx[d] = self.dx[d] if d < self.dim else 1.0
        return g

    def shape(self, ncomp):
        return (ncomp,) + tuple(reversed(self.nodes))

    def zeros(self, ncomp=None):
        return np.zeros(self.shape(self.dim if ncomp is None else ncomp))

    def coords(self):
        """Nodal coordinates, each of shape reversed(nodes)."""
        axes = [np.arange(self.nodes[d]) * self.dx[d] for d in range(self.dim)]
        mesh = np.meshgrid(*reversed(axes), indexing="ij")
        return list(reversed(mesh))

    def coarse(self):
        return Grid(tuple(v // 2 for v in self.n), tuple(self.n[d] * self.dx[d] for d in range(self.dim)))

    def fine(self):
        return Grid(tuple(v * 2 for v in self.n), tuple(self.n[d] * self.dx[d] for d in range(self.dim)))


def ncomp(dim, kind):
    return lib().or_ncomp(dim, kind)


def nfaces(dim):
    return 26 if dim == 3 else 8


def bc_arrays(dim, spec=None, default=(DISP, 0.0)):
    """Build (type[nfaces,3] int32, value[nfaces,3] f64) from {face: [(type, val) per comp]}."""
    nf = nfaces(dim)
    names = FACES3 if dim == 3 else FACES2
    t = np.full((nf, 3), default[0], dtype=np.int32)
    v = np.full((nf, 3), float(default[1]))
    for face, comps in (spec or {}).items():
        fi = names.index(face) if isinstance(face, str) else int(face)
        for c, tv in comps.items():
            t[fi, c] = tv[0]
            v[fi, c] = tv[1]
    return t, v


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def fapply(g, kind, C, bct, u, uniform=False):
    C, bct, u = _c(C), _c(bct, np.int32), _c(u)
    f = np.zeros_like(u)
    lib().or_fapply(ctypes.byref(g.c()), kind, _p(C), _p(bct), int(uniform), _p(u), _p(f))
    return f


def diagonal(g, kind, C, bct):
    C, bct = _c(C), _c(bct, np.int32)
    d = g.zeros()
    lib().or_diagonal(ctypes.byref(g.c()), kind, _p(C), _p(bct), _p(d))
    return d


def restrict(gc, fine):
    fine = _c(fine)
    nc = fine.shape[0]
    crse = gc.zeros(nc)
    lib().or_restrict(ctypes.byref(gc.c()), nc, _p(fine), _p(crse))
    return crse


def post(g, kind, C, u, what, voigt=False):
    """Strain (what 0), Stress (1) or Energy (2) of u (Elastic.cpp:399-616)."""
    nc = 1 if what == 2 else ((6 if g.dim == 3 else 3) if voigt else g.dim * g.dim)
    out = np.zeros(g.shape(nc))
    lib().or_post(ctypes.byref(g.c()), kind, _p(_c(C)), _p(_c(u)), what, int(bool(voigt)), _p(out))
    return out[0] if what == 2 else out


def newton_rhs(g, kind, C, bct, u, b, F0=None):
    """Newton<T>::prepareForSolve's right-hand side (Newton.H:30-140): b - div(DDW : (grad u - F0)) inside
    (central differences of dw), b - BC row(u, grad u, dw) on the boundary.  F0: (dim*dim, *nodes) or None."""
    out = np.zeros_like(_c(b))
    f0 = None if F0 is None else _c(F0)
    lib().or_newton_rhs(ctypes.byref(g.c()), kind, _p(_c(C)), None if f0 is None else _p(f0),
                        _p(_c(bct, np.int32)), _p(_c(u)), _p(_c(b)), _p(out))
    return out


def newton_w_dw(g, kind, C, u, F0=None, wfac=0.5):
    """Newton<T>::W and DW (Newton.H:233-333): (W per node, dw (dim*dim, *nodes))."""
    w = np.zeros(g.shape(1))[0]
    dw = np.zeros(g.shape(g.dim * g.dim))
    f0 = None if F0 is None else _c(F0)
    lib().or_newton_w_dw(ctypes.byref(g.c()), kind, _p(_c(C)), None if f0 is None else _p(f0), _p(_c(u)),
                         ctypes.c_double(wfac), _p(w), _p(dw))
    return w, dw


def _cell(eta, dim, i, j, k):
    """Cell values at cell index (i, j, k) of cells -1..n (one ghost layer): eta[c, k+1, j+1, i+1]."""
    if dim == 3:
        return eta[:, k + 1, j + 1, i + 1]
    return eta[:, j + 1, i + 1]


def degrade_material_nodes(g, eta_cells):
    """The node value PolymerDegradation::DegradeMaterial hands to DegradeModulus
    (PolymerDegradation.cpp:817-828): mul * (sum of the 2^dim cells around the node) in the reference's
    order, mul = 1/(2+2) in 2D and 1/(2+2+4) in 3D.  eta_cells: (ncomp, *(n + 2)) with one ghost layer.
    Returns (ncomp, *nodes).  Pure numpy (small grids)."""
    dim = g.dim
    I = np.arange(g.n[0] + 1)
    J = np.arange(g.n[1] + 1)
    if dim == 3:
        K = np.arange(g.n[2] + 1)
        e = lambda di, dj, dk: eta_cells[:, (K + dk + 1)[:, None, None], (J + dj + 1)[None, :, None],
                                         (I + di + 1)[None, None, :]]
        t = e(0, 0, 0) + e(-1, 0, 0)
        t = t + e(0, -1, 0) + e(-1, -1, 0)
        t = t + e(0, 0, -1) + e(-1, 0, -1) + e(0, -1, -1) + e(-1, -1, -1)
        return (1.0 / 8.0) * t
    e = lambda di, dj: eta_cells[:, (J + dj + 1)[:, None], (I + di + 1)[None, :]]
    t = e(0, 0) + e(-1, 0)
    t = t + e(0, -1) + e(-1, -1)
    return (1.0 / 4.0) * t


def pfm_eta_nodes(g, eta_cells):
    """PhaseFieldMicrostructure's node average (PFM.cpp:537): 0.25 (eta(i,j,k) + eta(i,j-1,k) + eta(i-1,j,k)
    + eta(i-1,j-1,k)), the x-y cells at cell layer k, also in 3D."""
    I = np.arange(g.n[0] + 1)
    J = np.arange(g.n[1] + 1)
    if g.dim == 3:
        K = np.arange(g.n[2] + 1)
        e = lambda di, dj: eta_cells[:, (K + 1)[:, None, None], (J + dj + 1)[None, :, None], (I + di + 1)[None, None, :]]
    else:
        e = lambda di, dj: eta_cells[:, (J + dj + 1)[:, None], (I + di + 1)[None, :]]
    return 0.25 * (e(0, 0) + e(0, -1) + e(-1, 0) + e(-1, -1))


def combine(ddw, F0, etas):
    """Affine::Cubic::Combine (Affine/Cubic.H:70-94) per node: ddw (nph, nc), F0 (nph, d, d) or None,
    etas (nph, *nodes).  Returns (C (nc, *nodes), F0 (d*d, *nodes) or None)."""
    etasum = np.zeros(etas.shape[1:])
    for n in range(etas.shape[0]):
        etasum = etasum + etas[n]
    C = np.zeros((ddw.shape[1],) + etas.shape[1:])
    F = None if F0 is None else np.zeros((F0.shape[1] * F0.shape[2],) + etas.shape[1:])
    for n in range(etas.shape[0]):
        w = etas[n] / etasum
        for q in range(ddw.shape[1]):
            C[q] = C[q] + ddw[n, q] * w
        if F is not None:
            for q in range(F.shape[0]):
                F[q] = F[q] + F0[n].reshape(-1)[q] * w
    return C, F


def normalize(g, diag, x):
    """Operator<Node>::normalize as the reference computes it (component 0 divided by diag
    component 0, once per component; Operator.cpp:164-195)."""
    out = _c(x).copy()
    lib().or_normalize(ctypes.byref(g.c()), _p(_c(diag)), _p(out))
    return out


def reflux(gc, crse_res, fine_origin, fine_res):
    """reflux (Operator.cpp:728-839): coarse residual nodes of the coarsened fine box take the
    full-weighting restriction of the fine residual (zero ghost values outside the box).

    crse_res: (ncomp, *coarse nodes reversed); fine_res: (ncomp, *box nodes reversed);
    fine_origin: fine-level node index (x, y[, z]) of the box's first node (even)."""
    out = _c(crse_res).copy()
    f = _c(fine_res)
    nc = f.shape[0]
    fn = list(reversed(f.shape[1:])) + [1]
    org = list(fine_origin) + [0]
    o = np.array(org[:3], dtype=np.int32)
    nb = np.array(fn[:3], dtype=np.int32)
    lib().or_reflux(ctypes.byref(gc.c()), nc, _p(out), _p(o), _p(nb), _p(f))
    return out


def interp_add(gf, crse, fine):
    crse = _c(crse)
    fine = _c(fine).copy()
    lib().or_interp_add(ctypes.byref(gf.c()), crse.shape[0], _p(crse), _p(fine))
    return fine


def fsmooth(g, kind, C, bct, diag, x, b):
    C, bct, diag, b = _c(C), _c(bct, np.int32), _c(diag), _c(b)
    x = _c(x).copy()
    lib().or_fsmooth(ctypes.byref(g.c()), kind, _p(C), _p(bct), _p(diag), _p(x), _p(b))
    return x


def bc_init(g, bct, bcv, rhs, homogeneous=False):
    bct, bcv = _c(bct, np.int32), _c(bcv)
    rhs = _c(rhs).copy()
    lib().or_bc_init(ctypes.byref(g.c()), _p(bct), _p(bcv), int(homogeneous), _p(rhs))
    return rhs


def cubic_define(C11, C12, C44, R, dim=3):
    R = _c(R).reshape(9)
    out = np.zeros(ncomp(dim, MM))
    lib().or_cubic_define_R(dim, C11, C12, C44, _p(R), _p(out))
    return out


def cubic_rotation(phi1, Phi, phi2):
    R = np.zeros(9)
    lib().or_cubic_rotation(phi1, Phi, phi2, _p(R))
    return R.reshape(3, 3)


def assemble(g, kind, C, bct):
    C, bct = _c(C), _c(bct, np.int32)
    n = g.dim * g.nnodes
    A = np.zeros((n, n))
    lib().or_assemble(ctypes.byref(g.c()), kind, _p(C), _p(bct), _p(A))
    return A


def mg_solve(g, kind, C, bct, b, u0=None, tol_rel=1e-10, tol_abs=0.0, max_iter=200, max_fmg_iter=10 ** 8,
             bottom="direct", levels=0, hist_cap=256, coarse_min=4):
    """Restated MLMG solve; returns (u, stats dict, residual history)."""
    C, bct, b = _c(C), _c(bct, np.int32), _c(b)
    u = np.zeros_like(b) if u0 is None else _c(u0).copy()
    stats = np.zeros(6)
    hist = np.zeros(hist_cap)
    rc = lib().or_mg_solve(ctypes.byref(g.c()), kind, _p(C), _p(bct), _p(b), _p(u), tol_rel, tol_abs, max_iter,
                           max_fmg_iter, 0 if bottom == "direct" else 1, levels, coarse_min, _p(stats), _p(hist),
                           hist_cap)
    it = int(stats[0])
    return u, {"converged": rc == 0, "iters": it, "res0": stats[1], "res": stats[2], "levels": int(stats[3]),
               "setup_s": stats[4], "iter_s": stats[5]}, hist[:it]


def set_threads(n):
    """OpenMP threads of the oracle's loops."""
    lib().or_set_threads(int(n))


def get_threads():
    return int(lib().or_get_threads())


def voronoi_ids(seeds_q, n, q):
    """Nearest-seed id per node of an (n[0], n[1], n[2])-cell grid, node i at i*q in seed units, ties to the
    lowest seed index; shape (nz+1, ny+1, nx+1).  Squared distances of these integer coordinates are below
    2^53, so the float64 distances of the k-d tree are exact and a tie is an exact equality; the 4 nearest
    candidates are compared and the lowest index among the equal nearest wins."""
    from scipy.spatial import cKDTree
    S = np.asarray(seeds_q, dtype=np.float64)
    tree = cKDTree(S)
    nx, ny, nz = (v + 1 for v in n)
    out = np.empty((nz, ny, nx), dtype=np.int32)
    J, I = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    kk = min(4, len(S))
    for k in range(nz):
        P = np.stack([I.ravel() * float(q), J.ravel() * float(q), np.full(I.size, k * float(q))], axis=1)
        d, idx = tree.query(P, k=kk, workers=-1)
        if kk == 1:
            out[k] = idx.reshape(ny, nx)
            continue
        best = idx[:, 0].copy()
        for c in range(1, kk):
            tie = (d[:, c] == d[:, 0]) & (idx[:, c] < best)
            best[tie] = idx[tie, c]
        out[k] = best.reshape(ny, nx)
    return out


def lame(E, nu):
    """(lambda, mu) from Young's modulus and Poisson ratio (Affine/Isotropic.H:67-74)."""
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = E / 2.0 / (1.0 + nu)
    return lam, mu
