"""Composite (AMR) solve on the GPU (ela_amr_*) against the composite restatement (oracle/amr.py)
and the reference's multi-level known-answer tests (UniaxialTest and TrigTest with 2 and 3 levels,
src/test.cc:128-150).

Tolerances: the composite residual is FP64 arithmetic in a different order from the oracle's
(1e-12 relative to the largest entry); solutions converged to 1e-10 agree with the oracle's sparse
direct solve to 1e-8 relative."""
import numpy as np
import pytest

import anisotropicfracture_amd as E
from tests.test_oracle import _uniaxial_bcs
from tests.test_oracle_amr import L_SHAPE, central, chop, trig_inputs, uniaxial_inputs

pytestmark = pytest.mark.gpu

ISO, DIAG = 0, 2


@pytest.fixture(scope="module")
def amr(oracle):
    from oracle import amr as A
    return A


def make(nest, kind, C, t, mode, coarse_min=4, uniform=False, values=None):
    a = E.ElasticAMR(nest.N[0], [nest.boxes[l] for l in range(1, nest.nlev)], kind, mode=mode,
                     coarse_min=coarse_min)
    for l in range(nest.nlev):
        assert a.shape(l) == nest.shape(l, nest.dim)
        assert a.origin(l) == nest.glo[l]
        a.SetModel(l, C[l])
    a.SetBC(t, values)
    a.SetUniform(uniform)
    a.prepareForSolve()
    return a


def random_iso(nest, rng):
    C = []
    for l in range(nest.nlev):
        c = np.zeros(nest.shape(l, 2))
        X = nest.coords(l)
        # smooth in space, so the levels describe the same material
        c[0] = 1.0 + 0.5 * np.sin(3 * X[0] + 1) * np.cos(2 * X[1]) + 0.3 * X[2]
        c[1] = 0.6 + 0.2 * np.cos(5 * X[2]) + 0.1 * X[0]
        C.append(c)
    return C


def mixed_bcs(oracle):
    return oracle.bc_arrays(3, {"ZLO": {0: (0, 0.0), 1: (0, 0.0), 2: (0, -0.01)},
                                "ZHI": {0: (1, 0.0), 1: (3, 0.0), 2: (0, 0.01)}}, default=(1, 0.0))


@pytest.mark.parametrize("n0,boxes", [
    (8, [((4,) * 3, (12,) * 3)]),
    (8, [((4,) * 3, (12,) * 3), ((12,) * 3, (20,) * 3)]),
    (8, [((0, 4, 4), (16, 12, 12))]),
    (12, L_SHAPE),
    (12, [chop((4, 4, 4), (20, 12, 12), 4), chop((12, 12, 12), (20, 20, 20), 4)]),
])
def test_composite_residual_matches_oracle(oracle, amr, n0, boxes):
    rng = np.random.default_rng(11)
    nest = amr.Nest((n0, n0, n0), boxes)
    C = random_iso(nest, rng)
    t, v = mixed_bcs(oracle)
    u = [rng.standard_normal(nest.shape(l, 3)) for l in range(nest.nlev)]
    b = [rng.standard_normal(nest.shape(l, 3)) for l in range(nest.nlev)]
    a = make(nest, ISO, C, t, E.SOLVE_FP64)
    rg, nrm = a.residual(u, b)
    ro, _ = nest.residual(ISO, C, t, b, u, uniform=False)
    scale = max(np.abs(x).max() for x in ro)
    for l in range(nest.nlev):
        print("level", l, "max diff", np.abs(rg[l] - ro[l]).max() / scale)
        assert np.abs(rg[l] - ro[l]).max() <= 1e-12 * scale
    assert abs(nrm - scale) <= 1e-12 * scale
    # the filled solution (slaves, ghosts, covered nodes) matches the oracle's fill
    uf = nest.fill(u)
    for l in range(nest.nlev):
        assert np.abs(a.download(a.U, l) - uf[l]).max() <= 1e-13 * max(np.abs(x).max() for x in uf)
    a.close()


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
@pytest.mark.parametrize("nlev", [2, 3])
def test_composite_solve_matches_oracle_direct(oracle, amr, mode, nlev):
    rng = np.random.default_rng(3)
    nest = amr.Nest((8, 8, 8), central(8, nlev))
    C = random_iso(nest, rng)
    t, v = mixed_bcs(oracle)
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(nest.nlev)]
    ref = nest.solve_direct(ISO, C, t, b)
    a = make(nest, ISO, C, t, mode)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=200)
    print("levels", nlev, "mode", mode, "iters", st["iters"], "res/res0", st["res"] / st["res0"])
    assert st["converged"] and st["res"] <= 1e-10 * st["res0"]
    scale = max(np.abs(x).max() for x in ref)
    for l in range(nest.nlev):
        m = nest.classify(l)["inB"]
        err = np.abs((u[l] - ref[l])[:, m]).max() / scale
        assert err <= 1e-8, (l, err)
    a.close()


@pytest.mark.parametrize("boxes", [
    [((16,) * 3, (48,) * 3)],
    [((16,) * 3, (48,) * 3), ((48,) * 3, (80,) * 3)],
    [((0, 16, 16), (64, 48, 48))],
])
def test_uniaxial_known_answer_composite_gpu(oracle, amr, boxes):
    """UniaxialTest at the reference's sizes (32^3 base; 2 levels, 3 levels, 2 non-centered levels):
    u_x = 0.1 x on every level."""
    nest = amr.Nest((32, 32, 32), boxes)
    C, t, b = uniaxial_inputs(nest, oracle)
    a = make(nest, ISO, C, t, E.SOLVE_MIXED, coarse_min=4)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=200)
    assert st["converged"]
    for l in range(nest.nlev):
        X = nest.coords(l)
        assert np.abs(u[l][0] - 0.1 * X[0]).max() < 1e-8
        assert np.abs(u[l][1:]).max() < 1e-8
    a.close()


@pytest.mark.parametrize("nlev", [2, 3])
def test_trig_known_answer_composite_gpu(oracle, amr, nlev):
    """TrigTest at the reference's sizes (32^3 base, Laplacian model): level-0 error below 2 %."""
    nest = amr.Nest((32, 32, 32), central(32, nlev))
    C, t, b, ex = trig_inputs(nest)
    a = make(nest, DIAG, C, t, E.SOLVE_MIXED)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=200)
    assert st["converged"]
    err = np.abs(u[0][0] - ex[0]).max() / np.abs(ex[0]).max()
    print("trig levels", nlev, "iters", st["iters"], "level-0 error", err)
    assert err < 0.02
    a.close()


def test_plain_fac_iteration_converges(oracle, amr):
    nest = amr.Nest((16, 16, 16), central(16, 2))
    C, t, b, ex = trig_inputs(nest)
    a = make(nest, DIAG, C, t, E.SOLVE_FP64)
    a.SetKrylov(0)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=400)
    print("plain FAC iters", st["iters"])
    assert st["converged"]
    ref = nest.solve_direct(DIAG, C, t, b)
    assert np.abs(u[0] - ref[0]).max() <= 1e-8 * np.abs(ref[0]).max()
    a.close()


def test_crack_nest_small_solution_is_a_composite_solution(oracle, amr):
    """The mode-I crack nest (3 levels, base 16^3): the GPU solution's composite residual, evaluated
    by the oracle, is at the solve's tolerance, in both precisions; mixed and FP64 agree."""
    from anisotropicfracture_amd import configs as K
    cfg = K.crack_nest(16)
    nest = amr.Nest(cfg["n"], cfg["boxes"])
    t, v = cfg["bc_type"], cfg["bc_value"]
    C = []
    for l in range(nest.nlev):
        X = nest.coords(l)
        C.append(K.crack_nest_model(cfg, X[0], X[1], X[2]))
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(nest.nlev)]
    r0, _ = nest.residual(ISO, C, t, b, [np.zeros_like(x) for x in b])
    r0 = max(np.abs(x).max() for x in r0)
    sols = {}
    for mode in (E.SOLVE_FP64, E.SOLVE_MIXED):
        a = make(nest, ISO, C, t, mode, values=v)
        for l in range(nest.nlev):
            assert np.array_equal(a.bc_init(l), b[l])  # BC::Init on the grown boxes
        u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=300)
        print("crack nest n=16 mode", mode, "iters", st["iters"], "res/res0", st["res"] / st["res0"])
        assert st["converged"]
        r, _ = nest.residual(ISO, C, t, b, u)
        assert max(np.abs(x).max() for x in r) <= 2e-10 * r0
        sols[mode] = u
        a.close()
    scale = max(np.abs(x).max() for x in sols[E.SOLVE_FP64])
    for l in range(nest.nlev):
        assert np.abs(sols[E.SOLVE_MIXED][l] - sols[E.SOLVE_FP64][l]).max() <= 1e-8 * scale


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
def test_composite_2d_matches_oracle(oracle, amr, mode):
    """2D nest of 3 levels (16^2 base): composite residual and solution against the oracle."""
    nest = amr.Nest((16, 16), [((8, 8), (24, 24)), ((24, 24), (40, 40))])
    t, v = oracle.bc_arrays(2, {"YLO": {0: (0, 0.0), 1: (0, 0.0)}, "YHI": {0: (1, 0.0), 1: (0, 0.01)},
                                "XLO_YLO": {0: (0, 0), 1: (0, 0)}, "XHI_YLO": {0: (0, 0), 1: (0, 0)}},
                            default=(1, 0.0))
    C = []
    for l in range(nest.nlev):
        X = nest.coords(l)
        c = np.zeros(nest.shape(l, 2))
        c[0] = 1.0 + 0.5 * np.sin(3 * X[0])
        c[1] = 0.7 + 0.2 * np.cos(2 * X[1])
        C.append(c)
    rng = np.random.default_rng(4)
    a = make(nest, ISO, C, t, mode, values=v)
    uu = [rng.standard_normal(nest.shape(l, 2)) for l in range(nest.nlev)]
    bb = [rng.standard_normal(nest.shape(l, 2)) for l in range(nest.nlev)]
    rg, _ = a.residual(uu, bb)
    ro, _ = nest.residual(ISO, C, t, bb, uu)
    scale = max(np.abs(x).max() for x in ro)
    for l in range(nest.nlev):
        assert np.abs(rg[l] - ro[l]).max() <= 1e-12 * scale
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 2))) for l in range(nest.nlev)]
    for l in range(nest.nlev):
        assert np.array_equal(a.bc_init(l), b[l])
    ref = nest.solve_direct(ISO, C, t, b)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=200)
    assert st["converged"]
    scale = max(np.abs(x).max() for x in ref)
    for l in range(nest.nlev):
        m = nest.classify(l)["inB"]
        assert np.abs((u[l] - ref[l])[:, m]).max() <= 1e-8 * scale
    a.close()


def test_bad_nesting_is_rejected():
    # 1 coarse cell inside its parent box (away from the domain boundary): the coarsened box grown by 2
    # coarse cells leaves the parent's union
    with pytest.raises(E.ElaError):
        E.ElasticAMR((8, 8, 8), [((4, 4, 4), (12, 12, 12)), ((10, 10, 10), (20, 20, 20))], ISO)
    # an L-shaped parent whose re-entrant corner is too close to the fine box
    with pytest.raises(E.ElaError):
        E.ElasticAMR((12, 12, 12), L_SHAPE[0:1] + [[((18, 18, 12), (22, 22, 16))]], ISO)
    with pytest.raises(E.ElaError):
        E.ElasticAMR((8, 8, 8), [((3, 4, 4), (12, 12, 12))], ISO)  # odd corner


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
def test_l_shaped_level_solve_matches_oracle_direct(oracle, amr, mode):
    """A refined level that is the union of three boxes in an L shape (its re-entrant corner gives
    coarse/fine faces of both orientations): the composite solution equals the oracle's sparse direct
    solve of the same composite system to 1e-8."""
    rng = np.random.default_rng(5)
    nest = amr.Nest((12, 12, 12), L_SHAPE)
    C = random_iso(nest, rng)
    t, v = mixed_bcs(oracle)
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(nest.nlev)]
    ref = nest.solve_direct(ISO, C, t, b)
    a = make(nest, ISO, C, t, mode)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=300)
    print("L shape mode", mode, "iters", st["iters"], "res/res0", st["res"] / st["res0"])
    assert st["converged"]
    scale = max(np.abs(x).max() for x in ref)
    for l in range(nest.nlev):
        m = nest.classify(l)["inB"]
        assert np.abs((u[l] - ref[l])[:, m]).max() <= 1e-8 * scale
    a.close()


def test_chopped_levels_solve_bit_for_bit_like_one_box(oracle, amr):
    """max_grid_size chopping changes the box list, not the union: the composite solve of a 3-level nest
    whose levels are chopped into boxes of at most 8 cells gives the bits of the unchopped nest."""
    from anisotropicfracture_amd import configs as K
    cfg = K.crack_nest(32)
    outs = []
    for size in (0, 8):
        boxes = [[b] if size == 0 else chop(b[0], b[1], size) for b in cfg["regions"]]
        nest = amr.Nest(cfg["n"], boxes)
        C = []
        for l in range(nest.nlev):
            X = nest.coords(l)
            C.append(K.crack_nest_model(cfg, X[0], X[1], X[2]))
        t, v = cfg["bc_type"], cfg["bc_value"]
        b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(nest.nlev)]
        a = make(nest, ISO, C, t, E.SOLVE_MIXED, values=v)
        u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=300)
        assert st["converged"]
        outs.append((u, st["iters"], sum(len(x) for x in boxes)))
        a.close()
    assert outs[1][2] > 3 * outs[0][2]
    assert outs[0][1] == outs[1][1]
    for x, y in zip(outs[0][0], outs[1][0]):
        assert np.array_equal(x, y)


def test_composite_solution_plot_file(oracle, amr, tmp_path):
    """The composite solution written as a nodal plot file (Integrator::WritePlotFile) reads back with the
    levels' boxes and the solution values of every box."""
    from tests.test_plotfile import read_plotfile
    nest = amr.Nest((12, 12, 12), L_SHAPE)
    C, t, b = uniaxial_inputs(nest, oracle)
    a = make(nest, ISO, C, t, E.SOLVE_MIXED)
    u, st = a.solve([np.zeros_like(x) for x in b], b, tol_rel=1e-10, max_iter=200)
    a.WritePlotFile(tmp_path / "plt00000node", {"disp": u}, time=0.0, step=0)
    pf = read_plotfile(str(tmp_path / "plt00000node"))
    assert pf["names"] == ["disp001", "disp002", "disp003"] and pf["finest"] == 1
    for l in range(nest.nlev):
        o = nest.glo[l]
        for (lo, hi, _), fab in zip(pf["levels"][l]["boxes"], pf["levels"][l]["fabs"]):
            want = u[l][:, lo[2] - o[2]:hi[2] - o[2] + 1, lo[1] - o[1]:hi[1] - o[1] + 1, lo[0] - o[0]:hi[0] - o[0] + 1]
            assert np.array_equal(fab, want)
    assert len(pf["levels"][1]["boxes"]) == 3
    a.close()
