"""ela_residual_noalloc (the allocation-free variant of ela_residual) against ela_residual, through the C ABI.

Both run the same kernels; the variant keeps its per-block maxima and norm word in the library's static
device array instead of the hierarchy's buffers.  The residual fields and norms must be the same bits, on
two calls in a row (the second on a smaller input, so the array holds the first call's maxima beyond the
second's), for the grain-form marching kernels, the plane kernels and both precisions.
"""
import ctypes

import numpy as np
import pytest

import anisotropicfracture_amd as ela
from anisotropicfracture_amd import configs as K
from tests import helpers as H

pytestmark = pytest.mark.gpu

ERR_BADARG = 1


def both(op, x, b, mglev, prec):
    L = ela.lib()
    fx = op.field(mglev, prec).upload(x)
    fb = op.field(mglev, prec).upload(b)
    r1, r2 = op.field(mglev, prec), op.field(mglev, prec)
    n1, n2 = ctypes.c_double(), ctypes.c_double()
    assert L.ela_residual(op.ptr, mglev, prec, fx.ptr, fb.ptr, r1.ptr, ctypes.byref(n1), None) == 0
    assert L.ela_residual_noalloc(op.ptr, mglev, prec, fx.ptr, fb.ptr, r2.ptr, ctypes.byref(n2), None) == 0, \
        L.ela_last_error()
    a, c = r1.download(), r2.download()
    for f in (fx, fb, r1, r2):
        f.free()
    return a, n1.value, c, n2.value


def polycrystal_op(n):
    cfg = K.polycrystal(n, grains=9, seed=31)
    op = ela.Elastic(cfg["n"], H.MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    op.SetBC(cfg["bc_type"], cfg["bc_value"])
    op.SetStiffnessMode(1)
    op.prepareForSolve()
    return op, cfg["n"]


def test_noalloc_residual_matches_residual_on_two_calls():
    rng = np.random.default_rng(17)
    # first call: a 128^3 grain-form polycrystal (LDS marching kernels on level 0, FP64 and FP32; a coarse level)
    op, n = polycrystal_op(128)
    shapes = [(0, ela.F64), (0, ela.F32), (1, ela.F64)]
    for mglev, prec in shapes:
        shp = op.shape(mglev, 3)
        x, b = rng.uniform(-1, 1, size=shp), rng.uniform(-1, 1, size=shp)
        a, na, c, nc = both(op, x, b, mglev, prec)
        assert np.array_equal(a, c), (mglev, prec)
        assert na == nc and na > 0, (mglev, prec, na, nc)
    op.close()
    # second call: a smaller input, per-node isotropic planes with mixed boundary rows
    m = (20, 16, 24)
    op2 = ela.Elastic(m, H.ISO, mode=ela.ALLOC_BOTH)
    op2.SetModel(H.random_stiffness(rng, m, H.ISO, 3))
    t, v = H.bc_mixed(3)
    op2.SetBC(t, v)
    op2.prepareForSolve()
    for prec in (ela.F64, ela.F32):
        x, b = H.random_field(rng, m, 3), H.random_field(rng, m, 3)
        a, na, c, nc = both(op2, x, b, 0, prec)
        assert np.array_equal(a, c) and na == nc and na > 0
        # the norm is the max |r| of the field (a stale maximum of the larger first call would show here)
        assert nc == np.abs(a).max()
    op2.close()


def test_noalloc_residual_rejects_levels_beyond_its_scratch():
    """A 2D level of 16385 x 2049 nodes needs 257 x 513 = 131841 block maxima, more than the 131072 the
    static array holds: a clear error, and ela_residual still serves it."""
    m = (16384, 2048)
    op = ela.Elastic(m, H.ISO, mode=ela.SOLVE_FP64, max_mg_levels=1)
    op.SetModel(np.ones((2, m[1] + 1, m[0] + 1)))
    fx = op.field(0, ela.F64).zero()
    fb = op.field(0, ela.F64).zero()
    nrm = ctypes.c_double()
    rc = ela.lib().ela_residual_noalloc(op.ptr, 0, ela.F64, fx.ptr, fb.ptr, None, ctypes.byref(nrm), None)
    assert rc == ERR_BADARG
    assert b"block maxima" in ela.lib().ela_last_error()
    assert ela.lib().ela_residual(op.ptr, 0, ela.F64, fx.ptr, fb.ptr, None, ctypes.byref(nrm), None) == 0
    assert nrm.value == 0.0
    fx.free()
    fb.free()
    op.close()
