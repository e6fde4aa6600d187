"""GPU parity: libela.so (gfx950 kernels, through the C ABI) against the CPU oracle.

Tolerances (FP64): operator outputs agree to 1e-12 relative to the scale of
the stencil terms (|C|max |u|max / h^2); FP32 to 2e-5 of that scale.  Solves
agree on the discrete solution to 1e-8 relative displacement.
"""
import numpy as np
import pytest

import anisotropicfracture_amd as ela
from anisotropicfracture_amd import configs as K
from tests import helpers as H

pytestmark = pytest.mark.gpu

ISO, MM, DIAG = H.ISO, H.MM, H.DIAG


def scale_of(C, u, h):
    return np.abs(C).max() * np.abs(u).max() / min(h) ** 2 + 1e-300


def make(n, kind, C, t, v=None, uniform=False, mode=ela.ALLOC_BOTH):
    op = ela.Elastic(n, kind, mode=mode)
    op.SetModel(C)
    op.SetBC(t, v)
    op.SetUniform(uniform)
    op.prepareForSolve()
    return op


CASES = [(2, ISO, (20, 13)), (2, MM, (17, 9)), (2, DIAG, (12, 12)),
         (3, ISO, (9, 7, 6)), (3, MM, (11, 6, 8)), (3, DIAG, (6, 6, 6)), (3, MM, (70, 5, 4)),
         # LDS-staged marching kernels: partial x tiles, partial y tiles, several z chunks
         (3, MM, (130, 9, 6)), (3, ISO, (40, 7, 70)), (3, DIAG, (20, 11, 37)), (3, MM, (33, 5, 67))]


@pytest.mark.parametrize("dim,kind,n", CASES)
@pytest.mark.parametrize("uniform", [False, True])
def test_fapply_parity(oracle, dim, kind, n, uniform):
    print("case", dim, kind, n, uniform, flush=True)
    rng = np.random.default_rng(hash((dim, kind, n)) % 2 ** 32)
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, dim)
    t, v = H.bc_mixed(dim)
    u = H.random_field(rng, n, dim)
    want = oracle.fapply(g, kind, C, t, u, uniform=uniform)
    op = make(n, kind, C, t, v, uniform)
    s = scale_of(C, u, g.dx)
    got = op.Fapply(u, prec=ela.F64)
    assert np.abs(got - want).max() <= 1e-12 * s
    got32 = op.Fapply(u, prec=ela.F32)
    assert np.abs(got32 - want).max() <= 2e-5 * s


@pytest.mark.parametrize("kind,n", [(MM, (9, 7, 6)), (MM, (40, 9, 36)), (ISO, (20, 5, 40)), (DIAG, (2, 3, 2))])
@pytest.mark.parametrize("prec", [ela.F64, ela.F32])
def test_residual_and_norm_parity(oracle, kind, n, prec):
    rng = np.random.default_rng(31)
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, 3)
    t, v = H.bc_mixed(3)
    op = make(n, kind, C, t, v)
    x, b = H.random_field(rng, n, 3), H.random_field(rng, n, 3)
    want = b - oracle.fapply(g, kind, C, t, x)
    r, nrm = op.residual(x, b, prec=prec)
    tol = (1e-12 if prec == ela.F64 else 2e-5) * scale_of(C, x, g.dx)
    assert np.abs(r - want).max() <= tol
    assert abs(nrm - np.abs(r).max()) <= 1e-300 + (0 if prec == ela.F64 else 1e-7 * nrm)
    assert abs(nrm - np.abs(want).max()) <= tol


@pytest.mark.parametrize("dim,kind,n", CASES[:6])
def test_diagonal_parity(oracle, dim, kind, n):
    rng = np.random.default_rng(5)
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, dim)
    t, v = H.bc_mixed(dim)
    op = make(n, kind, C, t, v)
    want = oracle.diagonal(g, kind, C, t)
    np.testing.assert_allclose(op.Diagonal(0, ela.F64), want, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(op.Diagonal(0, ela.F32), want, rtol=2e-7, atol=1e-30)


@pytest.mark.parametrize("n", [(128, 6, 4), (256, 4, 4)])
def test_prolongation_of_wide_levels(oracle, n):
    """Fine levels of 129+ nodes a row take the prolongation kernel with a thread
    per coarse node (the last coarse node of a row, column or plane has one fine
    node in that direction): FP64 against the oracle, FP32 against the oracle on
    the FP32-rounded inputs."""
    rng = np.random.default_rng(29)
    g = oracle.Grid(n)
    gc = g.coarse()
    op = ela.Elastic(n, MM, coarse_min=2)
    assert op.levels() >= 2
    f = H.random_field(rng, n, 3)
    c = H.random_field(rng, gc.n, 3)
    np.testing.assert_allclose(op.interpolation(f, c, 0), oracle.interp_add(g, c, f), rtol=1e-14, atol=1e-15)
    got = op.interpolation(f, c, 0, prec=ela.F32)
    want = oracle.interp_add(g, c.astype(np.float32).astype(np.float64), f.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dim,n", [(2, (16, 8)), (3, (16, 8, 8))])
def test_transfer_and_coarse_operator_parity(oracle, dim, n):
    rng = np.random.default_rng(17)
    kind = MM
    g = oracle.Grid(n)
    gc = g.coarse()
    C = H.random_stiffness(rng, n, kind, dim)
    t, v = H.bc_mixed(dim)
    op = make(n, kind, C, t, v)
    assert op.levels() >= 2
    f = H.random_field(rng, n, dim)
    np.testing.assert_allclose(op.restriction(f, 1), oracle.restrict(gc, f), rtol=1e-14, atol=1e-15)
    c = H.random_field(rng, gc.n, dim)
    np.testing.assert_allclose(op.interpolation(f, c, 0), oracle.interp_add(g, c, f), rtol=1e-14, atol=1e-15)
    # coarse operator uses the restricted stiffness (averageDownCoeffsSameAmrLevel)
    Cc = oracle.restrict(gc, C)
    uc = H.random_field(rng, gc.n, dim)
    want = oracle.fapply(gc, kind, Cc, t, uc)
    got = op.Fapply(uc, mglev=1)
    assert np.abs(got - want).max() <= 1e-12 * scale_of(Cc, uc, gc.dx)


def test_marching_jacobi_in_solver_matches_oracle_fsmooth(oracle):
    """One V-cycle-free check of the marching Jacobi: a 1-iteration solve with nu1 = nu2 = 0 is a pure
    coarse solve, so instead compare Fsmooth on a level large enough for the marching kernels."""
    rng = np.random.default_rng(12)
    n, kind = (40, 10, 34), MM
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, 3)
    t, v = H.bc_mixed(3)
    op = make(n, kind, C, t, v)
    x = H.random_field(rng, n, 3)
    want = oracle.fapply(g, kind, C, t, x)
    got = op.Fapply(x)
    assert np.abs(got - want).max() <= 1e-12 * scale_of(C, x, g.dx)


def test_fsmooth_parity(oracle):
    rng = np.random.default_rng(2)
    n, kind = (9, 8, 7), MM
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, 3)
    t, v = H.bc_mixed(3)
    op = make(n, kind, C, t, v)
    D = oracle.diagonal(g, kind, C, t)
    x, b = H.random_field(rng, n, 3), H.random_field(rng, n, 3)
    want = oracle.fsmooth(g, kind, C, t, D, x, b)
    got = op.Fsmooth(x, b, nsweeps=2)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_voronoi_stiffness_kernel(oracle):
    """Grain assignment bit-exact with the integer nearest-seed checker; stiffness values to 1e-15."""
    cfg = K.polycrystal(24, grains=17, seed=9)
    C_want, gid = H.voronoi_cubic_field(cfg, oracle.cubic_define)
    op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    op.SetBC(cfg["bc_type"], cfg["bc_value"])
    op.prepareForSolve()
    # read back through Fapply of unit-free probes is indirect; compare the operator instead
    rng = np.random.default_rng(1)
    u = H.random_field(rng, cfg["n"], 3)
    g = oracle.Grid(cfg["n"])
    want = oracle.fapply(g, MM, C_want, cfg["bc_type"], u)
    got = op.Fapply(u)
    assert np.abs(got - want).max() <= 1e-12 * scale_of(C_want, u, g.dx)


def _solve_case(oracle, cfg_n=16, grains=12):
    cfg = K.polycrystal(cfg_n, grains=grains, seed=3)
    C, _ = H.voronoi_cubic_field(cfg, oracle.cubic_define)
    g = oracle.Grid(cfg["n"])
    b = oracle.bc_init(g, cfg["bc_type"], cfg["bc_value"], g.zeros())
    return cfg, C, g, b


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_solve_matches_oracle_cubic(oracle, mode):
    print("solve mode", mode, flush=True)
    cfg, C, g, b = _solve_case(oracle)
    u_or, st_or, hist = oracle.mg_solve(g, MM, C, cfg["bc_type"], b, tol_rel=1e-12, max_iter=300)
    assert st_or["converged"]
    op = make(cfg["n"], MM, C, cfg["bc_type"], cfg["bc_value"])
    b_gpu = op.bc_init()
    np.testing.assert_array_equal(b_gpu, b)
    u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
    assert st["converged"]
    assert st["res"] / st["res0"] <= 1e-10
    err = np.abs(u - u_or).max() / np.abs(u_or).max()
    print("rel displacement error", err, "iters", st["iters"], "oracle iters", st_or["iters"])
    assert err <= 1e-8
    if mode == ela.SOLVE_FP64:
        # same algorithm: the iteration counts to 1e-10 agree within one cycle
        it_or = int(np.argmax(hist <= 1e-10 * st_or["res0"])) + 1
        assert abs(st["iters"] - it_or) <= 1


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_uniaxial_known_answer_gpu(mode):
    """UniaxialTest at the reference's 32^3: u_x = 0.1 x exactly (to solver tolerance)."""
    from tests.test_oracle import uniaxial_case
    n = 32
    C, t, v = uniaxial_case(n)
    op = make((n, n, n), ISO, C, t, v)
    b = op.bc_init()
    u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode)
    x = np.arange(n + 1) / n
    exact = 0.1 * np.broadcast_to(x, (n + 1,) * 3)
    assert np.abs(u[0] - exact).max() / 0.1 < 1e-8
    assert np.abs(u[1:]).max() < 1e-9


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_krylov_solve_matches_oracle_cubic(oracle, mode):
    """FGMRES around the MG cycle converges to the same discrete solution."""
    cfg, C, g, b = _solve_case(oracle)
    u_or, st_or, _ = oracle.mg_solve(g, MM, C, cfg["bc_type"], b, tol_rel=1e-12, max_iter=300)
    op = make(cfg["n"], MM, C, cfg["bc_type"], cfg["bc_value"])
    op.SetKrylov(16)
    u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
    assert st["converged"] and st["res"] / st["res0"] <= 1e-10
    assert np.abs(u - u_or).max() / np.abs(u_or).max() <= 1e-8


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
@pytest.mark.parametrize("nn", [32, 64])
def test_inclusion_2d_krylov_vs_direct(oracle, mode, nn):
    """The 2D inclusion (sharp 10x inclusion) at 32^2 and at the configuration's 64^2: the plain MG
    iteration of the reference diverges here; FGMRES(MG) reaches the discrete solution of a sparse direct
    solve.  The reference solution is the LU solve plus one FP64 refinement step: the LU solve alone stops
    at a relative residual of 2e-8 at 64^2 with an error of the same size, the refined one at the operator's
    FP64 floor (2.2e-9), where a second step changes it by 1e-15.  The GPU solves run to just above that
    floor.  Near the floor the iteration count is sensitive to the Krylov step's accuracy (a normalization
    taken from |w|^2 - |h|^2 instead of the stored vector raised it from 24 to 35 in FP64), so it is bounded
    here: measured 21 / 33 (FP64 / mixed) at 32^2 and 24 / 37 at 64^2."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    (nx, ny), C, t, v = K.inclusion_2d(nn)
    g = oracle.Grid((nx, ny))
    b = oracle.bc_init(g, t, v, g.zeros())
    lu = spl.splu(sp.csc_matrix(oracle.assemble(g, ISO, C, t)))
    u_ref = lu.solve(b.ravel()).reshape(b.shape)
    u_ref = u_ref + lu.solve((b - oracle.fapply(g, ISO, C, t, u_ref)).ravel()).reshape(b.shape)
    assert np.abs(b - oracle.fapply(g, ISO, C, t, u_ref)).max() / np.abs(b).max() < 3e-9
    op = make((nx, ny), ISO, C, t, v)
    op.SetKrylov(24)
    tol = 1e-9 if nn == 32 else 3e-9  # the floors: 5e-10 at 32^2, 2.2e-9 at 64^2
    u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=tol, tol_abs=0.0, mode=mode, max_iter=400)
    print("2D inclusion %d^2" % nn, st)
    assert st["converged"]
    assert np.abs(u - u_ref).max() / np.abs(u_ref).max() <= 1e-8
    bound = {(32, ela.SOLVE_FP64): 25, (32, ela.SOLVE_MIXED): 39, (64, ela.SOLVE_FP64): 28, (64, ela.SOLVE_MIXED): 43}
    assert st["iters"] <= bound[(nn, mode)], st


@pytest.mark.parametrize("n", [(24, 24, 24), (40, 20, 33)])
def test_grain_table_fapply_matches_planes_and_oracle(oracle, n):
    """Grain-table stiffness (gid + per-grain records) gives the same operator as the per-node planes."""
    cfg = K.polycrystal(8, grains=13, seed=11)
    rng = np.random.default_rng(4)
    cfg["n"] = n
    cfg["seeds"] = K.voronoi_seeds(rng, 13, n)
    C_want, _ = H.voronoi_cubic_field(cfg, oracle.cubic_define)
    op = ela.Elastic(n, MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    t, v = H.bc_mixed(3)
    op.SetBC(t, v)
    op.prepareForSolve()
    u = H.random_field(rng, n, 3)
    g = oracle.Grid(n)
    want = oracle.fapply(g, MM, C_want, t, u)
    s = scale_of(C_want, u, g.dx)
    planes64, planes32 = op.Fapply(u, prec=ela.F64), op.Fapply(u, prec=ela.F32)
    op.SetStiffnessMode(1)
    tab64, tab32 = op.Fapply(u, prec=ela.F64), op.Fapply(u, prec=ela.F32)
    assert np.abs(tab64 - want).max() <= 1e-12 * s
    assert np.abs(tab64 - planes64).max() <= 1e-13 * s
    assert np.abs(tab32 - want).max() <= 2e-5 * s
    r_tab, nrm_tab = op.residual(u, np.zeros_like(u), prec=ela.F64)
    assert np.abs(r_tab + want).max() <= 1e-12 * s


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
@pytest.mark.parametrize("krylov", [0, 16])
def test_grain_table_solve_matches_oracle(oracle, mode, krylov):
    cfg = K.polycrystal(16, grains=12, seed=3)
    C, _ = H.voronoi_cubic_field(cfg, oracle.cubic_define)
    g = oracle.Grid(cfg["n"])
    b = oracle.bc_init(g, cfg["bc_type"], cfg["bc_value"], g.zeros())
    u_or, _, _ = oracle.mg_solve(g, MM, C, cfg["bc_type"], b, tol_rel=1e-12, max_iter=300)
    op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    op.SetBC(cfg["bc_type"], cfg["bc_value"])
    op.SetStiffnessMode(1)
    op.SetKrylov(krylov)
    op.prepareForSolve()
    u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
    assert st["converged"] and st["res"] / st["res0"] <= 1e-10
    assert np.abs(u - u_or).max() / np.abs(u_or).max() <= 1e-8


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
@pytest.mark.parametrize("krylov", [0, 16])
def test_grain_marching_kernels_give_the_node_kernel_bits(mode, krylov):
    """Level 0 in grain form through the LDS z-marching kernels (one and two rows per thread, one row with
    two planes loaded ahead; the tiles' interior Jacobi diagonal from the grains' records, and from the
    stored field with OPT_GRAIN_DIAG 0) and through the one-thread-per-node kernels: the same arithmetic,
    so the same solve and the same Fapply bit for bit (a 40^3 box has partial tiles in x, y, z)."""
    cfg = K.polycrystal(40, grains=14, seed=5)
    u1 = H.random_field(np.random.default_rng(8), cfg["n"], 3)
    sols, iters, fa = [], [], []
    for march, rows, pf, gd in ((0, 1, 1, 1), (1, 1, 1, 1), (1, 2, 1, 1), (1, 1, 2, 1), (1, 1, 1, 0)):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.SetKrylov(krylov)
        op.set_option(ela.OPT_MARCH, march)
        op.set_option(ela.OPT_ROWS, rows)
        op.set_option(ela.OPT_PREFETCH, pf)
        op.set_option(ela.OPT_GRAIN_DIAG, gd)
        op.prepareForSolve()
        fa.append((op.Fapply(u1, prec=ela.F64), op.Fapply(u1, prec=ela.F32)))
        b = op.bc_init()
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
        assert st["converged"]
        sols.append(u)
        iters.append(st["iters"])
        op.close()
    for v in (1, 2, 3, 4):
        assert iters[0] == iters[v]
        assert np.array_equal(sols[0], sols[v])
        assert np.array_equal(fa[0][0], fa[v][0]) and np.array_equal(fa[0][1], fa[v][1])


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_per_grain_diagonal_keeps_the_bits_on_coarse_marching_levels(mode):
    """A 256^3 polycrystal marches on level 1 (129^3) too, with its own spacing: the tiles' interior Jacobi
    diagonal per grain and level (OPT_GRAIN_DIAG 1) gives the bits of the stored diagonal field (0),
    cycle for cycle, over ten FGMRES iterations."""
    cfg = K.polycrystal(256, grains=64, seed=2)
    out = []
    for gd in (0, 1):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.set_option(ela.OPT_GRAIN_DIAG, gd)
        op.prepareForSolve()
        b = op.bc_init()
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-30, tol_abs=0.0, mode=mode, max_iter=10,
                                   raise_on_fail=False)
        out.append((u, st["res"]))
        op.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_plane_mask_keeps_the_bits(mode):
    """Per-node-plane stiffness: skipping the derivative of C where the records at a node's two
    neighbours are bitwise equal, with the constant-C nodes launched first (OPT_MASK 2: node lists,
    OPT_MASK 3: regrouped inside each block; every level), against never skipping (OPT_MASK 0).  The skipped terms are exact zeros and every node is computed
    once in either order, so Fapply, the residual and whole solves agree bit for bit (numpy equality,
    which also takes -0.0 == 0.0).  Few large grains, so whole wavefronts do skip."""
    cfg = K.polycrystal(48, grains=5, seed=9)
    out = {}
    for m in (0, 2, 3):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(0)
        op.set_option(ela.OPT_MASK, m)
        op.prepareForSolve()
        u = H.random_field(np.random.default_rng(3), cfg["n"], 3)
        b = op.bc_init()
        f64, f32 = op.Fapply(u, prec=ela.F64), op.Fapply(u, prec=ela.F32)
        r, nrm = op.residual(u, b, prec=ela.F64)
        x, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
        assert st["converged"]
        out[m] = (f64, f32, r, nrm, x, st["iters"])
        op.close()
    for m in (2, 3):
        for a, c in zip(out[0], out[m]):
            assert np.array_equal(a, c)


@pytest.mark.parametrize("n", [(37, 20, 45), (40, 17, 33)])
@pytest.mark.parametrize("mask", [0, 1])
def test_plane_marching_kernels_give_the_node_kernel_bits(oracle, n, mask):
    """Isotropic per-node planes (every node its own record: the derivative of C everywhere, except in a
    constant half) with the mixed boundary table: level 0 through the LDS z-marching kernels (one and two
    planes loaded ahead) and through the one-thread-per-node kernels (OPT_MARCH 0) give the same Fapply
    (FP64, FP32), residual and norm bit for bit; Fapply agrees with the oracle to 1e-12 of the stencil
    scale.  With the mask (OPT_MASK 2) the marching tiles read each node's same-direction byte."""
    rng = np.random.default_rng(41)
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, ISO, 3)
    C[:, : (n[0] + 1) // 2] = C[:, :1, :1, :1]  # a constant half: its nodes skip the derivative with the mask
    t, v = H.bc_mixed(3)
    u, b = H.random_field(rng, n, 3), H.random_field(rng, n, 3)
    out = []
    for march, pf in ((0, 1), (1, 1), (1, 2)):
        op = ela.Elastic(n, ISO, mode=ela.ALLOC_BOTH)
        op.SetModel(C)
        op.SetBC(t, v)
        op.set_option(ela.OPT_MARCH, march)
        op.set_option(ela.OPT_PREFETCH, pf)
        op.set_option(ela.OPT_MASK, 2 * mask)
        op.prepareForSolve()
        res = [op.Fapply(u, prec=ela.F64), op.Fapply(u, prec=ela.F32)]
        for prec in (ela.F64, ela.F32):
            r, nrm = op.residual(u, b, prec=prec)
            res += [r, np.array([nrm])]
        op.close()
        out.append(res)
    want = oracle.fapply(g, ISO, C, t, u)
    assert np.abs(out[1][0] - want).max() <= 1e-12 * scale_of(C, u, g.dx)
    names = ("Fapply F64", "Fapply F32", "r F64", "|r| F64", "r F32", "|r| F32")
    for v_ in (1, 2):
        for name, a, c in zip(names, out[0], out[v_]):
            bad = np.argwhere(a != c)
            assert len(bad) == 0, (v_, name, len(bad), bad[:4].tolist(), float(np.abs(a - c).max()))


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_plane_marching_solve_gives_the_node_kernel_bits(mode):
    """A smoothly varying isotropic stiffness under uniaxial z loading, FGMRES(6): the solve through the
    marching kernels on level 0 is the solve through the node kernels, bit for bit."""
    n = (40, 36, 44)
    x = [np.linspace(0.0, 1.0, m + 1) for m in n]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    C = np.empty((2,) + X.shape)
    C[0] = 1.2 * (1.0 + 0.5 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y))
    C[1] = 0.8 * (1.0 + 0.4 * np.cos(2 * np.pi * Z) * np.sin(np.pi * X))
    C = np.moveaxis(C, (1, 2, 3), (3, 2, 1)).copy()  # (comp, z, y, x)
    cfg = K.polycrystal(8)  # its uniaxial z boundary table
    sols = []
    for march in (0, 1):
        op = ela.Elastic(n, ISO, mode=ela.ALLOC_BOTH)
        op.SetModel(C)
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetKrylov(6)
        op.set_option(ela.OPT_MARCH, march)
        op.prepareForSolve()
        b = op.bc_init()
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
        assert st["converged"]
        sols.append((u, st["iters"]))
        op.close()
    assert sols[0][1] == sols[1][1]
    assert np.array_equal(sols[0][0], sols[1][0])


@pytest.mark.parametrize("n", [(40, 36, 34), (128, 128, 128)])
def test_grain_form_coarse_operator_is_the_injected_one(oracle, n):
    """A grain-form hierarchy coarsens by injection: level 1's operator is the oracle's operator with the
    stiffness of the fine nodes (2I, 2J, 2K), i.e. every coarse node keeps its grain's record.  A level
    of 65+ nodes a side below the marching size runs the grain-table node kernels, a smaller one the
    per-node planes."""
    cfg = K.polycrystal(8, grains=9, seed=21)
    cfg["n"] = n
    cfg["seeds"] = K.voronoi_seeds(np.random.default_rng(21), 9, n)
    C, _ = H.voronoi_cubic_field(cfg, oracle.cubic_define)
    op = ela.Elastic(n, MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    t, v = H.bc_mixed(3)
    op.SetBC(t, v)
    op.SetStiffnessMode(1)
    op.prepareForSolve()
    gc = oracle.Grid(n).coarse()
    Cc = np.ascontiguousarray(C[:, ::2, ::2, ::2])
    uc = H.random_field(np.random.default_rng(5), gc.n, 3)
    want = oracle.fapply(gc, MM, Cc, t, uc)
    s = scale_of(Cc, uc, gc.dx)
    assert np.abs(op.Fapply(uc, mglev=1, prec=ela.F64) - want).max() <= 1e-12 * s
    assert np.abs(op.Fapply(uc, mglev=1, prec=ela.F32) - want).max() <= 2e-5 * s
    op.close()


def test_coarse_marching_levels_give_the_plane_kernel_bits():
    """On a 256^3 grain-form hierarchy, level 1 (129^3) runs the LDS marching kernels (with the tile heights
    the autotuner picks for levels 0 and 1); with OPT_MARCH 0 it runs the per-node kernels on the same
    injected records.  Both evaluate the same expressions, so level 1's operator and the whole mixed solve
    agree bit for bit (numpy equality)."""
    cfg = K.polycrystal(256, grains=16, seed=4)
    u1 = H.random_field(np.random.default_rng(8), (128, 128, 128), 3)
    out = []
    for march in (0, 1):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.SetKrylov(6)
        op.set_option(ela.OPT_MARCH, march)
        op.prepareForSolve()
        if march:
            op.autotune(nrep=1)
        f64, f32 = op.Fapply(u1, mglev=1, prec=ela.F64), op.Fapply(u1, mglev=1, prec=ela.F32)
        b = op.field(0, ela.F64).upload(op.bc_init())
        u = op.field(0, ela.F64)
        u.zero()
        rc, st = op.solve_device(b, u, 1e-10, 0.0, ela.SOLVE_MIXED, 100)
        assert rc == 0 and st["converged"]
        out.append((f64, f32, u.download(), st["iters"]))
        op.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def test_autotune_picks_a_tile_shape_without_changing_results():
    cfg = K.polycrystal(40, grains=10, seed=6)
    sols = []
    for tune in (False, True):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.prepareForSolve()
        if tune:
            times = op.autotune(nrep=3)
            assert len(times) == 18 and all(t > 0 for t in times.values())  # 12 one-plane, 6 two-plane shapes
        b = op.bc_init()
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=ela.SOLVE_MIXED, max_iter=300)
        assert st["converged"]
        sols.append(u)
        op.close()
    assert np.array_equal(sols[0], sols[1])


@pytest.mark.parametrize("case", ["polycrystal", "iso_vcycle"])
def test_graph_replayed_cycles_keep_the_bits(case):
    """OPT_GRAPH 1 replays the cycles' launches on levels >= 1 from captured HIP graphs (one per buffer
    state: the F-cycle's exchanges alternate two).  Several solves in a row (each replays the graphs many
    times between eager level-0 work), a changed option and a new set-up in between (new captures) and both
    precisions must give the stream launches' bits."""
    out = []
    for graph in (0, 1):
        if case == "polycrystal":
            cfg = K.polycrystal(48, grains=7, seed=12)
            op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
            op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
            op.SetBC(cfg["bc_type"], cfg["bc_value"])
            op.SetStiffnessMode(1)
            b = None
        else:
            n = (16, 16, 32)
            op = make(n, ISO, H.random_stiffness(np.random.default_rng(3), n, ISO, 3), *H.bc_mixed(3))
            op.SetCycle(ela.CYCLE_V, 2, 2)
            b = H.random_field(np.random.default_rng(4), n, 3)
        op.SetKrylov(6)
        op.set_option(ela.OPT_GRAPH, graph)
        op.prepareForSolve()
        if b is None:
            b = op.bc_init()
        res = []
        for rep in range(3):
            if rep == 2:
                op.set_option(ela.OPT_KZ, 8)
                op.prepareForSolve()
            for mode in (ela.SOLVE_MIXED, ela.SOLVE_FP64):
                u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode,
                                             max_iter=300)
                assert st["converged"]
                res.append((u, st["iters"]))
        out.append(res)
        op.close()
    for (ua, ia), (ub, ib) in zip(out[0], out[1]):
        assert ia == ib
        assert np.array_equal(ua, ub)


def test_graph_eviction_keeps_the_bits():
    """More graphs than the hierarchy keeps (8): every tile height in turn gives each solve's launches new
    keys (4 a precision: two kinds times two buffer states), so captured graphs are evicted and destroyed
    while later ones replay, and a height seen before is captured again after its eviction.  An evicted
    graph is destroyed only after the stream has drained.  Every solve must keep the stream launches' bits."""
    cfg = K.polycrystal(48, grains=7, seed=12)
    out = []
    for graph in (0, 1):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.SetKrylov(6)
        op.set_option(ela.OPT_GRAPH, graph)
        op.prepareForSolve()
        b = op.bc_init()
        res = []
        for kz in (16, 8, 4, 32, 16, 8):
            op.set_option(ela.OPT_KZ, kz)
            for mode in (ela.SOLVE_MIXED, ela.SOLVE_FP64):
                u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode,
                                             max_iter=300)
                assert st["converged"]
                res.append((u, st["iters"]))
        out.append(res)
        op.close()
    for (ua, ia), (ub, ib) in zip(out[0], out[1]):
        assert ia == ib
        assert np.array_equal(ua, ub)


@pytest.mark.parametrize("mode", [ela.SOLVE_FP64, ela.SOLVE_MIXED])
def test_folded_last_sweep_gives_the_separate_addition_bits(mode):
    """OPT_FOLD 1 (default): the F-cycle's last level-0 sweep of a Krylov step writes the FP64 vector
    (its value + the held correction) itself; OPT_FOLD 0 stores the sweep's value and adds in a separate
    kernel.  The FGMRES iterates, hence the iteration count and solution, must be the same bits."""
    cfg = K.polycrystal(40, grains=6, seed=21)
    out = []
    for fold in (0, 1):
        op = ela.Elastic(cfg["n"], MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetStiffnessMode(1)
        op.SetKrylov(6)
        op.set_option(ela.OPT_FOLD, fold)
        op.prepareForSolve()
        b = op.bc_init()
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
        assert st["converged"]
        out.append((u, st["iters"]))
        op.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("n,kind", [((32, 32, 32), MM), ((56, 64, 72), ISO), ((48, 40), MM)])
def test_gauss_jordan_one_launch_a_column_gives_the_launch_pair_bits(n, kind):
    """OPT_GJ 1 (default): ela_prepare inverts the dense coarsest operator by Gauss-Jordan in one launch a
    column (the pivot search in every block, the row swaps as a permutation, the pivot row divided where
    it is read); OPT_GJ 0 launches the pivot search and the elimination separately.  Same pivots and
    operations, so FP64 and mixed solves give the same iterations and solution bits.  Coarsest levels of
    375 (5^3 nodes), 2160 (8 x 9 x 10 nodes) and 2D unknowns, with traction and Neumann faces."""
    dim = len(n)
    C = H.random_stiffness(np.random.default_rng(sum(n)), n, kind, dim)
    t, v = H.bc_mixed(dim)
    for mode in (ela.SOLVE_FP64, ela.SOLVE_MIXED):
        out = []
        for gj in (0, 1):
            op = make(n, kind, C, t, v)
            op.set_option(ela.OPT_GJ, gj)
            op.SetKrylov(6)
            op.prepareForSolve()
            b = op.bc_init()
            u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=300)
            assert st["converged"], (n, gj, st)
            out.append((u, st["iters"]))
            op.close()
        print(n, "mode", mode, "iters", out[0][1], out[1][1])
        assert out[0][1] == out[1][1]
        assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("n,mode", [((36, 36, 28), ela.SOLVE_MIXED)])
def test_sizes_without_a_dense_coarsest_level_solve_with_a_bicgstab_bottom(oracle, n, mode):
    """A size whose coarsest level is too large for the dense inverse (36 x 36 x 28 -> 9 x 9 x 7 cells,
    2400 unknowns) solves with the BiCGStab bottom (AMReX's MLCGSolver with MLMG's bottom tolerance
    1e-4), plain MG and FGMRES, to the solution of the oracle's MG with its BiCGStab bottom, to 1e-8
    relative (the oracle's solve takes about 20 s of CPU, so one case)."""
    rng = np.random.default_rng(sum(n))
    dim = len(n)
    kind = MM if dim == 3 else ISO
    C = H.random_stiffness(rng, n, kind, dim)
    t, v = K.uniaxial_z_bcs(0.01) if dim == 3 else K.inclusion_2d(8)[2:]
    op = make(n, kind, C, t, v)
    g = oracle.Grid(n)
    b = oracle.bc_init(g, t, v, g.zeros())
    u_or, st_or, _ = oracle.mg_solve(g, kind, C, t, b, tol_rel=1e-11, max_iter=120, bottom="bicgstab")
    assert st_or["res"] <= 5e-11 * st_or["res0"], st_or
    for krylov in (0, 8):
        op.SetKrylov(krylov)
        u, st = ela.Linear(op).solve(np.zeros_like(b), b, tol_rel=1e-10, tol_abs=0.0, mode=mode, max_iter=400)
        print(n, "krylov", krylov, "iters", st["iters"], "levels", op.levels())
        assert st["converged"]
        assert np.abs(u - u_or).max() <= 1e-8 * np.abs(u_or).max()
    op.close()
