"""The composite (AMR) restatement (oracle/amr.py) pinned by the reference's own multi-level
known-answer tests and by invariants of the transfers.

* UniaxialTest with 2 and 3 levels and with "2 non-centered levels" (Elastic_UniaxialTest.cpp:13-206,
  src/test.cc:140-150): the affine field u_x = 0.1 x solves the composite system exactly.
* TrigTest with 2 and 3 levels (Elastic_TrigTest.cpp:17-140, src/test.cc:128-134): level-0 error of
  the composite solution below the reference's 2 %, and second-order convergence.
* RefluxTest (Elastic_RefluxTest.cpp:14-108): affine fields leave zero residual after reflux.
* The reflux weights are a partition of unity at every coarse node (a constant fine residual,
  ghosts included, restricts to the constant); interpolation reproduces affine fields exactly.
* The direct solve (sparse matrix form) and the residual function (array form) are built by
  separate code; the direct solution's residual is at round-off.
"""
import numpy as np
import pytest

from tests.test_oracle import _uniaxial_bcs

ISO, DIAG = 0, 2


@pytest.fixture(scope="module")
def amr(oracle):
    from oracle import amr as A
    return A


def central(n, nlev):
    """The reference test geometry (Test/Operator/Elastic.cpp:56-89, Grid::XYZ): each level the
    refinement of the middle half of the one below."""
    boxes = []
    lo, hi = n // 2, 3 * n // 2
    for _ in range(1, nlev):
        boxes.append(((lo,) * 3, (hi,) * 3))
        lo, hi = 2 * lo + (hi - lo) // 2, 2 * hi - (hi - lo) // 2
    return boxes


def uniaxial_inputs(nest, O):
    t, v = _uniaxial_bcs()
    C = [np.zeros(nest.shape(l, 2)) for l in range(nest.nlev)]
    for c in C:
        c[0], c[1] = 6.0, 2.6
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(nest.nlev)]
    return C, t, b


def trig_inputs(nest):
    t = np.zeros((26, 3), dtype=np.int32)
    C = [np.ones(nest.shape(l, 9)) for l in range(nest.nlev)]
    b, ex = [], []
    for l in range(nest.nlev):
        X = nest.coords(l)
        tr = np.sin(np.pi * X[0]) * np.sin(np.pi * X[1]) * np.sin(np.pi * X[2])
        bb = np.zeros(nest.shape(l, 3))
        bb[0] = tr
        b.append(nest.bc_init(l, t, np.zeros((26, 3)), bb))
        ex.append(-tr / (3 * np.pi ** 2))
    return C, t, b, ex


def test_central_geometry_matches_reference_generator():
    # Generate(): level 1 = refine(grow(domain, -n/4)), level 2 = refine(grow(that, -n/4))
    assert central(32, 3) == [((16,) * 3, (48,) * 3), ((48,) * 3, (80,) * 3)]


@pytest.mark.parametrize("boxes", [
    [((4,) * 3, (12,) * 3)],
    [((4,) * 3, (12,) * 3), ((12,) * 3, (20,) * 3)],
    [((0, 4, 4), (16, 12, 12))],  # Grid::YZ, "2 non-centered levels"
])
def test_uniaxial_composite_exact(oracle, amr, boxes):
    nest = amr.Nest((8, 8, 8), boxes)
    C, t, b = uniaxial_inputs(nest, oracle)
    u = nest.solve_direct(ISO, C, t, b)
    for l in range(nest.nlev):
        X = nest.coords(l)
        assert np.abs(u[l][0] - 0.1 * X[0]).max() < 1e-10
        assert np.abs(u[l][1:]).max() < 1e-10
    r, _ = nest.residual(ISO, C, t, b, u)
    assert max(np.abs(x).max() for x in r) < 1e-10


def test_trig_composite_known_answer_and_order(oracle, amr):
    errs = []
    for n in (8, 16):
        nest = amr.Nest((n,) * 3, central(n, 2))
        C, t, b, ex = trig_inputs(nest)
        u = nest.solve_direct(DIAG, C, t, b)
        e0 = np.abs(u[0][0] - ex[0]).max() / np.abs(ex[0]).max()
        e1 = np.abs(u[1][0] - ex[1]).max() / np.abs(ex[0]).max()
        assert e0 < 0.02  # the reference's tolerance
        errs.append((e0, e1))
    # second order: halving h divides the error by about 4
    assert errs[0][0] / errs[1][0] > 3.5
    assert errs[0][1] / errs[1][1] > 3.0


def test_trig_composite_three_levels(oracle, amr):
    nest = amr.Nest((8,) * 3, central(8, 3))
    C, t, b, ex = trig_inputs(nest)
    u = nest.solve_direct(DIAG, C, t, b)
    assert np.abs(u[0][0] - ex[0]).max() / np.abs(ex[0]).max() < 0.02


def test_reflux_partition_of_unity_and_affine_interpolation(oracle, amr):
    nest = amr.Nest((8, 8, 8), [((4, 0, 4), (12, 16, 12))])
    crse = np.zeros(nest.shape(0, 3))
    fine = np.full(nest.shape(1, 3), 2.5)
    out = nest.reflux(1, crse, fine)
    sl = tuple(slice(v // 2, w // 2 + 1) for v, w in zip(reversed(nest.lo[1]), reversed(nest.hi[1])))
    assert np.allclose(out[(slice(None),) + sl], 2.5, rtol=0, atol=1e-15)
    X = nest.coords(0)
    a = np.stack([0.3 * X[0] - 0.2 * X[1] + 0.7 * X[2], X[1], -X[2]])
    p = nest.prolong(1, a)
    Xf = nest.coords(1)
    ref = np.stack([0.3 * Xf[0] - 0.2 * Xf[1] + 0.7 * Xf[2], Xf[1], -Xf[2]])
    assert np.abs(p - ref).max() < 1e-14


def test_reflux_test_affine_fields(oracle, amr):
    """RefluxTest: affine solutions with a uniform isotropic model give zero residual on both levels,
    and the refluxed coarse residual stays zero (< 1e-12)."""
    nest = amr.Nest((8, 8, 8), [((4,) * 3, (12,) * 3)])
    t, _ = oracle.bc_arrays(3)
    C = [np.zeros(nest.shape(l, 2)) for l in range(2)]
    for c in C:
        c[0], c[1] = 6.0, 2.6
    for nvec in ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [1, 1, 0]):
        u = []
        for l in range(2):
            X = nest.coords(l)
            s = nvec[0] * X[0] + nvec[1] * X[1] + nvec[2] * X[2]
            ul = np.zeros(nest.shape(l, 3))
            ul[1] = s
            u.append(ul)
        b = [oracle.fapply(nest.grid(l), ISO, C[l], nest.bct_array(l, t), u[l]) for l in range(2)]
        r, _ = nest.residual(ISO, C, t, b, u)
        assert max(np.abs(x).max() for x in r) < 1e-12


def test_direct_solution_has_zero_composite_residual(oracle, amr):
    rng = np.random.default_rng(5)
    nest = amr.Nest((8, 8, 8), [((4,) * 3, (12,) * 3)])
    t, v = oracle.bc_arrays(3, {"ZLO": {0: (0, 0.0), 1: (0, 0.0), 2: (0, -0.01)},
                                "ZHI": {0: (1, 0.0), 1: (1, 0.0), 2: (0, 0.01)}}, default=(1, 0.0))
    C = []
    for l in range(2):
        c = np.zeros(nest.shape(l, 2))
        c[0] = 1.0 + rng.random(c[0].shape)
        c[1] = 0.5 + rng.random(c[1].shape)
        C.append(c)
    b = [nest.bc_init(l, t, v, np.zeros(nest.shape(l, 3))) for l in range(2)]
    u = nest.solve_direct(ISO, C, t, b)
    r, _ = nest.residual(ISO, C, t, b, u)
    r0, _ = nest.residual(ISO, C, t, b, [np.zeros_like(x) for x in b])
    assert max(np.abs(x).max() for x in r) < 1e-10 * max(np.abs(x).max() for x in r0)


# ---------------------------------------------------------------- levels made of several boxes
L_SHAPE = [[((4, 4, 4), (12, 12, 12)), ((12, 4, 4), (20, 12, 12)), ((4, 12, 4), (12, 20, 12))]]


def chop(lo, hi, size):
    """A box cut into boxes of at most `size` cells per direction (AMReX's max_grid_size chopping)."""
    axes = [list(range(lo[d], hi[d], size)) + [hi[d]] for d in range(len(lo))]
    out = []
    for corner in np.ndindex(*[len(a) - 1 for a in axes]):
        out.append((tuple(axes[d][corner[d]] for d in range(len(lo))),
                    tuple(axes[d][corner[d] + 1] for d in range(len(lo)))))
    return out


def classes_by_points(nest, l):
    """Node classes restated with real coordinates and closed/open box containment: a node is in the
    union when it lies in a closed box; a touching cell (its centre at the node +- h/2) counts when its
    centre is inside the domain; it is in the union (under level l+1) when the centre lies in a box (a
    level-(l+1) box scaled by 1/2)."""
    dim = nest.dim
    X = [x.ravel() for x in nest.gidx(l)]
    N = nest.N[l]
    inbox = lambda p, boxes, s: any(all(lo[d] * s <= p[d] <= hi[d] * s for d in range(dim)) for lo, hi in boxes)
    out = {k: np.zeros(X[0].size, bool) for k in ("inB", "slave", "covered", "iface")}
    for q in range(X[0].size):
        p = [float(X[d][q]) for d in range(dim)]
        out["inB"][q] = inbox(p, nest.boxes[l], 1.0)
        cells = [[p[d] + o[d] - 0.5 for d in range(dim)] for o in np.ndindex(*(2,) * dim)]
        cells = [c for c in cells if all(0 < c[d] < N[d] for d in range(dim))]
        inu = [inbox(c, nest.boxes[l], 1.0) for c in cells]
        out["slave"][q] = l > 0 and out["inB"][q] and not all(inu)
        if l + 1 < nest.nlev:
            und = [inbox(c, nest.boxes[l + 1], 0.5) for c in cells]
            out["covered"][q] = all(und) and out["inB"][q]
            out["iface"][q] = any(und) and not all(und)
    shp = nest.shape(l, 1)[1:]
    return {k: v.reshape(shp) for k, v in out.items()}


@pytest.mark.parametrize("boxes", [L_SHAPE, [chop((4, 4, 4), (20, 12, 12), 4)],
                                   [[((4, 4, 4), (12, 12, 12)), ((12, 4, 4), (20, 12, 12))], [((12, 12, 12), (20, 20, 20))]]])
def test_union_classes_match_a_point_containment_restatement(amr, boxes):
    nest = amr.Nest((12, 12, 12), boxes)
    for l in range(nest.nlev):
        want = classes_by_points(nest, l)
        got = nest.classify(l)
        for k in want:
            assert np.array_equal(got[k], want[k]), (l, k)


def test_chopped_levels_give_the_same_composite_problem(oracle, amr):
    """A region chopped into boxes of at most 4 cells (max_grid_size) is the same union: the same
    classes, the same composite residual and the same direct solution as the region as one box."""
    one = amr.Nest((8, 8, 8), [((4, 4, 4), (12, 12, 12))])
    many = amr.Nest((8, 8, 8), [chop((4, 4, 4), (12, 12, 12), 4)])
    assert len(many.boxes[1]) == 8
    for l in range(2):
        a, b = one.classify(l), many.classify(l)
        for k in a:
            assert np.array_equal(a[k], b[k])
    C, t, bb = uniaxial_inputs(one, oracle)
    rng = np.random.default_rng(4)
    u = [rng.standard_normal(one.shape(l, 3)) for l in range(2)]
    r1, _ = one.residual(ISO, C, t, bb, u)
    r2, _ = many.residual(ISO, C, t, bb, u)
    for x, y in zip(r1, r2):
        assert np.array_equal(x, y)


def test_uniaxial_composite_exact_on_an_l_shaped_level(oracle, amr):
    """UniaxialTest (Elastic_UniaxialTest.cpp:13-206) with a level that is the union of three boxes in an
    L shape (its re-entrant corner makes coarse/fine faces of both orientations): u_x = 0.1 x is still
    the composite solution, and the direct solve's composite residual is at round-off."""
    nest = amr.Nest((12, 12, 12), L_SHAPE)
    C, t, b = uniaxial_inputs(nest, oracle)
    u = nest.solve_direct(ISO, C, t, b)
    for l in range(nest.nlev):
        X = nest.coords(l)
        m = nest.classify(l)["inB"]
        assert np.abs((u[l][0] - 0.1 * X[0])[m]).max() < 1e-10
        assert np.abs(u[l][1:][:, m]).max() < 1e-10
    r, _ = nest.residual(ISO, C, t, b, u)
    assert max(np.abs(x).max() for x in r) < 1e-10
