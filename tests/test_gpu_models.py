"""GPU parity of the model-field assembly and of Newton's W / DW (through the C ABI).

DegradeModulus (Degradable/Isotropic.H:74-104), DegradeMaterial
(PolymerDegradation.cpp:793-833), Combine with PFM's node average
(PFM.cpp:529-543, Affine/Cubic.H:70-94) write the level-0 stiffness on the
device; each is compared with the oracle through the operator it defines
(Fapply, 1e-12 of the stencil scale) and, for Combine, through the F0 planes
it returns (1e-15 relative).  W / DW against the oracle to 1e-12.  The error
paths return the reference's abort conditions as status codes.
"""
import numpy as np
import pytest

import anisotropicfracture_amd as ela
from tests import helpers as H

pytestmark = pytest.mark.gpu

ISO, MM = H.ISO, H.MM
ERR_BADARG, ERR_NAN = 1, 5


def scale_of(C, u, h):
    return np.abs(C).max() * np.abs(u).max() / min(h) ** 2 + 1e-300


def check_operator(oracle, op, n, kind, C_want, bct, rng):
    g = oracle.Grid(n)
    u = H.random_field(rng, n, len(n))
    f = op.Fapply(u)
    f_or = oracle.fapply(g, kind, C_want, bct, u)
    assert np.abs(f - f_or).max() <= 1e-12 * scale_of(C_want, u, g.dx)


def iso_op(n, bct):
    op = ela.Elastic(n, ISO, mode=ela.SOLVE_FP64)
    op.SetBC(bct)
    return op


@pytest.mark.parametrize("n", [(12, 10), (8, 6, 10)])
def test_degrade_modulus_one_and_two_eta(oracle, n):
    rng = np.random.default_rng(3)
    bct, _ = H.bc_mixed(len(n))
    lam0, mu0 = 121.15, 80.77
    shp = tuple(v + 1 for v in reversed(n))
    op = iso_op(n, bct)
    eta = rng.uniform(0, 1, size=shp)
    op.SetModelDegradedIsotropic(lam0, mu0, eta)
    check_operator(oracle, op, n, ISO, np.stack([lam0 * (1.0 - eta), mu0 * (1.0 - eta)]), bct, rng)
    eta2 = rng.uniform(0, 1, size=(2,) + shp)
    op.SetModelDegradedIsotropic(lam0, mu0, eta2)
    check_operator(oracle, op, n, ISO, np.stack([lam0 * (1.0 - eta2[0]), mu0 * (1.0 - eta2[1])]), bct, rng)
    op.close()


def test_degrade_modulus_rejects_what_the_reference_aborts_on(oracle):
    """NaN anywhere -> ELA_ERR_NAN; an infinite eta0 -> ELA_ERR_NAN ("Infs in eta"); a value outside
    [0, 1], an infinite eta1 included (the reference's Inf test reads eta0 only) -> ELA_ERR_BADARG.
    The model stays as it was."""
    n = (8, 8, 8)
    rng = np.random.default_rng(4)
    bct, _ = H.bc_mixed(3)
    op = iso_op(n, bct)
    shp = tuple(v + 1 for v in reversed(n))
    eta = rng.uniform(0, 1, size=(2,) + shp)
    op.SetModelDegradedIsotropic(2.0, 1.0, eta)
    u = H.random_field(rng, n, 3)
    f0 = op.Fapply(u)
    cases = [((0, 3, 2, 1), np.nan, ERR_NAN), ((1, 1, 1, 1), np.nan, ERR_NAN), ((0, 0, 0, 0), np.inf, ERR_NAN),
             ((1, 5, 5, 5), np.inf, ERR_BADARG), ((0, 2, 2, 2), 1.5, ERR_BADARG), ((1, 4, 4, 4), -0.1, ERR_BADARG)]
    for idx, val, code in cases:
        bad = eta.copy()
        bad[idx] = val
        with pytest.raises(ela.ElaError) as ei:
            op.SetModelDegradedIsotropic(2.0, 1.0, bad)
        assert ei.value.code == code, (idx, val)
        assert np.array_equal(op.Fapply(u), f0)
    op.close()


@pytest.mark.parametrize("n", [(10, 12), (8, 6, 10)])
def test_degrade_material_matches_oracle(oracle, n):
    """Cell damage with a ghost layer, averaged onto the nodes; two components: the last one sets the
    moduli, as the reference's per-component scalar DegradeModulus calls leave it."""
    rng = np.random.default_rng(7)
    g = oracle.Grid(n)
    bct, _ = H.bc_mixed(len(n))
    cells = rng.uniform(0, 1, size=(2,) + tuple(v + 2 for v in reversed(n)))
    op = iso_op(n, bct)
    op.DegradeMaterial(3.0, 2.0, cells)
    e = oracle.degrade_material_nodes(g, cells)[1]
    check_operator(oracle, op, n, ISO, np.stack([3.0 * (1.0 - e), 2.0 * (1.0 - e)]), bct, rng)
    bad = cells.copy()
    bad[0] = bad[0] + 1.0  # component 0 out of range (its node means lie in [1, 2])
    with pytest.raises(ela.ElaError) as ei:
        op.DegradeMaterial(3.0, 2.0, bad)
    assert ei.value.code == ERR_BADARG
    op.close()


@pytest.mark.parametrize("n", [(10, 8, 12), (12, 10)])
def test_combine_matches_oracle(oracle, n):
    rng = np.random.default_rng(9)
    dim = len(n)
    g = oracle.Grid(n)
    bct, _ = H.bc_mixed(dim)
    nph = 3
    if dim == 3:
        ddw = np.stack([oracle.cubic_define(1.68, 1.21, 0.75, H.random_rotation(rng)) for _ in range(nph)])
    else:
        ddw = rng.uniform(0.5, 2.0, size=(nph, 6))
    F0 = rng.normal(size=(nph, 3, 3)) * 1e-3
    cells = rng.uniform(0.05, 1.0, size=(nph,) + tuple(v + 2 for v in reversed(n)))
    op = ela.Elastic(n, MM, mode=ela.SOLVE_FP64)
    op.SetBC(bct)
    F = op.CombineCubic(ddw, cells, F0, want_F0=True)
    C_or, F_or = oracle.combine(ddw, F0[:, :dim, :dim], oracle.pfm_eta_nodes(g, cells))
    check_operator(oracle, op, n, MM, C_or, bct, rng)
    np.testing.assert_allclose(F, F_or, rtol=1e-15, atol=1e-18)
    zero = cells.copy()
    zero[..., 1:3, 1:3] = 0.0  # the weights of some nodes sum to zero
    with pytest.raises(ela.ElaError) as ei:
        op.CombineCubic(ddw, zero)
    assert ei.value.code == ERR_NAN
    op.close()


@pytest.mark.parametrize("dim,kind,wfac,with_F0", [(3, MM, 0.5, False), (3, MM, 0.5, True), (2, ISO, 1.0, True),
                                                  (3, ISO, 0.5, False)])
def test_w_dw_match_oracle(oracle, dim, kind, wfac, with_F0):
    rng = np.random.default_rng(11 + dim)
    n = (10, 8, 12)[:dim]
    g = oracle.Grid(n)
    C = H.random_stiffness(rng, n, kind, dim)
    u = H.random_field(rng, n, dim)
    F0 = rng.normal(size=(dim * dim,) + tuple(v + 1 for v in reversed(n))) if with_F0 else None
    op = ela.Elastic(n, kind, mode=ela.SOLVE_FP64)
    op.SetModel(C)
    w, dw = op.W_DW(u, F0, wfac)
    w_or, dw_or = oracle.newton_w_dw(g, kind, C, u, F0, wfac)
    sc = np.abs(dw_or).max()
    assert np.abs(dw - dw_or).max() <= 1e-13 * sc
    assert np.abs(w - w_or).max() <= 1e-12 * np.abs(w_or).max()
    op.close()
