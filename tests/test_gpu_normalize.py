"""normalize (Operator.cpp:164-195) on the GPU against the oracle."""
import numpy as np
import pytest

import anisotropicfracture_amd as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_normalize_matches_oracle(oracle, dim):
    n = (6, 5, 4)[:dim]
    g = oracle.Grid(n)
    rng = np.random.default_rng(dim)
    lam, mu = oracle.lame(1.0, 0.3)
    C = np.stack([np.full(g.shape(1)[1:], lam), np.full(g.shape(1)[1:], mu)]) * (1 + 0.2 * rng.random(g.shape(2)))
    bct = np.zeros((26 if dim == 3 else 8, 3), np.int32)
    op = E.Elastic(n, E.ISO)
    op.SetModel(C)
    op.SetBC(bct)
    op.prepareForSolve()
    x = rng.standard_normal(g.shape(dim))
    got = op.normalize(x)
    want = oracle.normalize(g, oracle.diagonal(g, oracle.ISO, C, bct), x)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    op.close()
