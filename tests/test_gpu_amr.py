"""AMR level transfers on the GPU (ela_reflux, ela_average_down) against the oracle."""
import numpy as np
import pytest

import anisotropicfracture_amd as E

pytestmark = pytest.mark.gpu

CASES = [
    (3, 8, (4, 2, 6), (8, 10, 6)),   # inside the domain
    (3, 6, (0, 0, 4), (12, 6, 8)),   # touching the low faces and the x-high face
    (2, 10, (6, 0), (14, 20)),       # 2D, touching y faces
]


@pytest.mark.parametrize("dim,nc,origin,cells", CASES)
@pytest.mark.parametrize("prec", [E.F64, E.F32])
def test_reflux_matches_oracle(oracle, dim, nc, origin, cells, prec):
    rng = np.random.default_rng(dim * 100 + nc)
    crse = E.Elastic((nc,) * dim, E.ISO)
    box = E.Elastic(cells, E.ISO)
    gc = oracle.Grid((nc,) * dim)
    r0 = rng.standard_normal(gc.shape(dim))
    r1 = rng.standard_normal((dim,) + tuple(reversed([c + 1 for c in cells])))
    if prec == E.F32:
        r0, r1 = r0.astype(np.float32).astype(np.float64), r1.astype(np.float32).astype(np.float64)
    cf = crse.field(0, prec, dim).upload(r0)
    ff = box.field(0, prec, dim).upload(r1)
    E.reflux(cf, ff, origin)
    got = cf.download()
    want = oracle.reflux(gc, r0, origin, r1)
    tol = 0 if prec == E.F64 else 2e-6
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max()) + (4e-16 if prec == E.F64 else 0)
    crse.close()
    box.close()


def test_average_down_injects_coincident_nodes():
    rng = np.random.default_rng(5)
    crse = E.Elastic((8, 8, 8), E.ISO)
    box = E.Elastic((6, 4, 8), E.ISO)
    origin = (2, 8, 4)
    c0 = rng.standard_normal((3, 9, 9, 9))
    f = rng.standard_normal((3, 9, 5, 7))
    cf = crse.field(0, E.F64, 3).upload(c0)
    E.average_down(cf, box.field(0, E.F64, 3).upload(f), origin)
    got = cf.download()
    want = c0.copy()
    want[:, 2:7, 4:7, 1:5] = f[:, ::2, ::2, ::2]
    assert np.array_equal(got, want)
    crse.close()
    box.close()
