"""Newton<T>'s right-hand side (ela_newton_rhs) and the affine Newton solve on the GPU, against the oracle
(oracle.newton_rhs, restating src/Solver/Nonlocal/Newton.H:30-140).

Tolerances: the rhs kernel does the oracle's arithmetic (one-sided wall gradient, contraction, central
differences in the reference's order) in FP64, within 1e-12 of the largest entry; solves converged to
1e-10 agree with the oracle's to 1e-8 relative."""
import numpy as np
import pytest

import anisotropicfracture_amd as E
from tests import helpers as H

pytestmark = pytest.mark.gpu

ISO, MM = 0, 1


def mixed_bcs(oracle):
    return oracle.bc_arrays(3, {"ZLO": {0: (0, 0.0), 1: (0, 0.0), 2: (0, 0.0)},
                                "ZHI": {0: (1, 0.0), 1: (3, 0.0), 2: (0, 0.01)}}, default=(1, 0.0))


@pytest.mark.parametrize("kind", [ISO, MM])
def test_newton_rhs_matches_oracle(oracle, kind):
    rng = np.random.default_rng(7 + kind)
    n = (9, 7, 8)
    g = oracle.Grid(n)
    nc = 2 if kind == ISO else 21
    C = np.zeros(g.shape(nc))
    if kind == ISO:
        C[0], C[1] = 1.0 + rng.random(C[0].shape), 0.5 + rng.random(C[1].shape)
    else:
        for q in range(nc):
            C[q] = H.cubic_table_numpy(1.68, 1.21, 0.75, H.random_rotation(rng))[q] * (1 + 0.2 * rng.random(C[q].shape))
    t, v = mixed_bcs(oracle)
    u = rng.standard_normal(g.shape(3)) * 0.01
    b = rng.standard_normal(g.shape(3))
    F0 = rng.standard_normal(g.shape(9)) * 0.01
    op = E.Elastic(n, kind, mode=E.SOLVE_FP64)
    op.SetModel(C)
    op.SetBC(t, v)
    for f0 in (None, F0):
        got = op.newton_rhs(u, b, f0)
        ref = oracle.newton_rhs(g, kind, C, t, u, b, f0)
        scale = np.abs(ref).max()
        assert np.abs(got - ref).max() <= 1e-12 * scale
    zero = op.newton_rhs(np.zeros_like(u), b, None)
    assert np.array_equal(zero, b)  # u = 0, F0 = 0: exactly b
    op.close()


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
def test_affine_newton_inclusion_matches_oracle(oracle, mode):
    """An eigenstrain inclusion (Affine::Isotropic with F0 = 0.01 I inside a sphere, 0 outside): one Newton
    step from u = 0 on the GPU against the oracle's solve of A du = newton_rhs."""
    n = (16, 16, 16)
    g = oracle.Grid(n)
    X = g.coords()
    inside = (X[0] - 0.5) ** 2 + (X[1] - 0.5) ** 2 + (X[2] - 0.5) ** 2 < 0.25 ** 2
    C = np.zeros(g.shape(2))
    C[0], C[1] = 6.0, 2.6
    F0 = np.zeros(g.shape(9))
    for a in range(3):
        F0[4 * a][inside] = 0.01
    t, v = mixed_bcs(oracle)
    b = oracle.bc_init(g, t, v, g.zeros())
    rhs = oracle.newton_rhs(g, ISO, C, t, g.zeros(), b, F0)
    ref, st, _ = oracle.mg_solve(g, ISO, C, t, rhs, tol_rel=1e-12, tol_abs=0.0)
    assert st["converged"]
    op = E.Elastic(n, ISO, mode=E.ALLOC_BOTH)
    op.SetModel(C)
    op.SetBC(t, v)
    op.SetUniform(False)
    op.prepareForSolve()
    u, st = E.Newton(op).solve(np.zeros(g.shape(3)), b, F0, tol_rel=1e-10, tol_abs=0.0, mode=mode)
    assert st["converged"]
    assert np.abs(u - ref).max() <= 1e-8 * np.abs(ref).max()
    # the eigenstrain is what loads the body: the interior of the inclusion expands
    assert np.abs(u).max() > 1e-3
    op.close()


def test_linear_newton_is_one_linear_solve(oracle):
    """F0 = 0 and u0 = 0: Newton's step solves A du = b, bit for bit the Linear solve."""
    n = (16, 16, 16)
    g = oracle.Grid(n)
    C = np.zeros(g.shape(2))
    C[0], C[1] = 6.0, 2.6
    t, v = mixed_bcs(oracle)
    b = oracle.bc_init(g, t, v, g.zeros())
    op = E.Elastic(n, ISO, mode=E.SOLVE_FP64)
    op.SetModel(C)
    op.SetBC(t, v)
    op.prepareForSolve()
    u1, _ = E.Newton(op).solve(np.zeros(g.shape(3)), b, None, tol_rel=1e-10, tol_abs=0.0)
    u2, _ = E.Linear(op).solve(np.zeros(g.shape(3)), b, tol_rel=1e-10, tol_abs=0.0)
    assert np.array_equal(u1, u2)
    op.close()
