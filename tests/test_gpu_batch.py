"""Batched independent solves (ela_batch / ela_rve) against the oracle.

Each member is a cubic Voronoi polycrystal on [0,1]^3 with Displacement rows on
every face and the affine boundary displacement u = G x (the fixed-corner
"periodic-like" RVE loading); some members carry a crack degradation factor.
The checker is the CPU oracle on per-node stiffness planes rebuilt from the
member's grain ids and the closed-form cubic table.
"""
import numpy as np
import pytest

import anisotropicfracture_amd as E
import anisotropicfracture_amd.configs as K
from helpers import cubic_table_numpy

pytestmark = pytest.mark.gpu

def problem(seed, n, ngrains, crack=False):
    rng = np.random.default_rng(seed)
    seeds = K.voronoi_seeds(rng, ngrains, (n, n, n))
    R = K.haar_rotations(rng, ngrains)
    a = rng.standard_normal(3)
    a /= np.linalg.norm(a)
    G = 0.01 * np.outer(a, a)
    cr = None
    if crack:
        nrm = rng.standard_normal(3)
        nrm /= np.linalg.norm(nrm)
        tv = np.cross(nrm, rng.standard_normal(3))
        tv /= np.linalg.norm(tv)
        # soft crack: g = (1 - eta)^2 + 1e-3
        cr = (rng.random(3) * 0.5 + 0.25, nrm, tv, 2.0 / n, (1.0, -2.0, 1.0 + 1e-3))
    return {"n": n, "seeds": seeds, "R": R, "G": G, "crack": cr}


def make(batch, p):
    r = batch.rve(p["n"], max_grains=len(p["seeds"]), with_factor=p["crack"] is not None)
    r.SetModelVoronoiCubic(p["seeds"], K.VORONOI_Q, p["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
    if p["crack"] is not None:
        c0, nrm, tv, width, g = p["crack"]
        r.SetCrack(c0, nrm, tv, width, *g)
    r.SetRhsAffine(p["G"])
    return r


def planes(r, p):
    """Per-node stiffness planes of the member, rebuilt on the host."""
    gid = r.grain_ids().astype(np.int64)
    tab = np.stack([cubic_table_numpy(K.C11, K.C12, K.C44, Ri) for Ri in p["R"]])
    C = np.moveaxis(tab[gid], -1, 0).copy()
    if p["crack"] is not None:
        C *= r.download(E.RVE.FACTOR)[None]
    return C


def all_disp():
    return np.zeros((26, 3), dtype=np.int32)


def oracle_solution(oracle, r, p, C):
    n = p["n"]
    b = r.download(E.RVE.B)
    u, st, _ = oracle.mg_solve(oracle.Grid((n, n, n)), oracle.MM, C, all_disp(), b, tol_rel=1e-14, coarse_min=2)
    assert st["res"] <= 1e-12 * st["res0"], st
    return u


def test_member_inputs_match_closed_forms():
    b = E.Batch(E.SOLVE_MIXED)
    p = problem(1, 16, 12, crack=True)
    r = make(b, p)
    n = p["n"]
    gid = r.grain_ids()
    assert np.array_equal(gid.astype(np.int64), K.voronoi_grain_ids(p["seeds"], (n, n, n)))
    # degradation factor: eta from the distance to the half-plane
    x = np.arange(n + 1) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    P = np.stack([X, Y, Z])
    c0, nrm, tv, w, (a, bb, cc) = p["crack"]
    dn = np.tensordot(nrm, P - np.asarray(c0)[:, None, None, None], 1)
    dt = np.tensordot(tv, P - np.asarray(c0)[:, None, None, None], 1)
    eta = 0.5 * (1 - np.tanh(np.sqrt(dn ** 2 + np.maximum(dt, 0) ** 2) / w))
    g = (a * eta + bb) * eta + cc
    np.testing.assert_allclose(r.download(E.RVE.FACTOR), g, rtol=1e-13, atol=1e-15)
    # right-hand side: G x on every boundary node (all rows Displacement), 0 inside
    rhs = r.download(E.RVE.B)
    want = np.tensordot(p["G"], P, 1)
    bnd = np.zeros((n + 1,) * 3, bool)
    bnd[[0, -1]] = True
    bnd[:, [0, -1]] = True
    bnd[:, :, [0, -1]] = True
    np.testing.assert_allclose(rhs[:, bnd], want[:, bnd], rtol=1e-14, atol=1e-18)
    assert not rhs[:, ~bnd].any()
    r.close()
    b.close()


@pytest.mark.parametrize("prec,tol", [(E.F64, 1e-12), (E.F32, 2e-5)])
def test_batched_fapply_matches_oracle(oracle, prec, tol):
    b = E.Batch(E.SOLVE_MIXED)
    probs = [problem(11, 16, 8), problem(12, 12, 5, crack=True), problem(13, 10, 3)]
    rs = [make(b, p) for p in probs]
    # boundary rows of every kind: traction faces, displacement faces, a mix
    bcts = [all_disp(), all_disp(), all_disp()]
    bcts[0][:] = K.TRAC
    bcts[2][[0, 1, 3, 4]] = K.TRAC
    bcts[2][6:, 0] = K.TRAC
    rng = np.random.default_rng(3)
    us = []
    for r, bct in zip(rs, bcts):
        r.SetBC(bct)
        u = rng.standard_normal(r.shape)
        r.upload(E.RVE.U, u)
        us.append(u)
    b.fapply(rs, prec)  # one launch over the three members
    for r, p, u, bct in zip(rs, probs, us, bcts):
        n = p["n"]
        f = r.download(E.RVE.R)
        want = oracle.fapply(oracle.Grid((n, n, n)), oracle.MM, planes(r, p), bct, u)
        scale = np.abs(want).max()
        assert np.abs(f - want).max() <= tol * scale, (n, np.abs(f - want).max() / scale)
    for r in rs:
        r.close()
    b.close()


def test_batched_solve_matches_oracle(oracle):
    b = E.Batch(E.SOLVE_MIXED)
    probs = [problem(21, 16, 10), problem(22, 24, 16), problem(23, 20, 6, crack=True), problem(24, 12, 4)]
    rs = [make(b, p) for p in probs]
    b.prepare(rs)
    sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
    for r, p, st in zip(rs, probs, sts):
        print("n=%d levels=%d cycles=%d res/res0=%.2e" % (p["n"], r.levels(), st["iters"], st["res"] / st["res0"]))
        assert st["converged"] and st["res"] <= 1e-10 * st["res0"]
        u = r.download(E.RVE.U)
        ref = oracle_solution(oracle, r, p, planes(r, p))
        err = np.abs(u - ref).max() / np.abs(ref).max()
        assert err <= 1e-8, (p["n"], err)
    b.close()


def test_batch_member_runs_the_iteration_of_the_grain_form_single_box():
    """A batch member keeps grain-form stiffness on every level (coarse grain ids injected at set-up), the
    operators of a grain-form single-box hierarchy with the same coarsening (coarse_min 2): in FP64 both take
    the same cycles and reach the same solution to 1e-12 relative (the plain FP64 baseline of the batched
    RVE workload relies on it).  Displacement on every face, the member's affine boundary values."""
    for seed, n, ng in ((91, 32, 12), (92, 24, 7)):
        p = problem(seed, n, ng)
        b = E.Batch(E.SOLVE_FP64)
        r = make(b, p)
        b.prepare([r])
        st_b = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
        ub = r.download(E.RVE.U)
        bh = r.download(E.RVE.B)
        b.close()
        op = E.Elastic((n, n, n), E.MM, mode=E.SOLVE_FP64, coarse_min=2)
        op.SetBC(np.zeros((26, 3), np.int32), np.zeros((26, 3)))
        op.SetModelVoronoiCubic(p["seeds"], K.VORONOI_Q, p["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
        op.SetStiffnessMode(1)
        op.SetKrylov(0)
        op.prepareForSolve()
        us, st_s = E.Linear(op).solve(np.zeros_like(bh), bh, tol_rel=1e-10, tol_abs=0.0, mode=E.SOLVE_FP64, max_iter=80)
        op.close()
        print("n=%d batch cycles %d single-box cycles %d" % (n, st_b["iters"], st_s["iters"]))
        assert st_b["converged"] and st_s["converged"]
        assert st_b["iters"] == st_s["iters"]
        assert np.abs(ub - us).max() <= 1e-12 * np.abs(us).max()


def test_fp64_batch_matches_mixed():
    sol = {}
    for mode in (E.SOLVE_FP64, E.SOLVE_MIXED):
        b = E.Batch(mode)
        p = problem(31, 16, 8)
        r = make(b, p)
        b.prepare([r])
        st = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
        assert st["converged"]
        sol[mode] = r.download(E.RVE.U)
        b.close()
    ref = sol[E.SOLVE_FP64]
    assert np.abs(sol[E.SOLVE_MIXED] - ref).max() <= 1e-8 * np.abs(ref).max()


def test_member_result_is_the_same_alone_and_in_a_mix():
    target = problem(41, 16, 9, crack=True)
    others = [problem(42, 32, 12), problem(43, 12, 5), problem(44, 20, 7, crack=True)]
    b = E.Batch(E.SOLVE_MIXED)
    r = make(b, target)
    b.prepare([r])
    st_alone = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
    alone = r.download(E.RVE.U)
    b.close()

    b = E.Batch(E.SOLVE_MIXED)
    rs = [make(b, p) for p in others[:2]] + [make(b, target)] + [make(b, others[2])]
    b.prepare(rs)
    sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
    mixed = rs[2].download(E.RVE.U)
    assert sts[2]["iters"] == st_alone["iters"]
    assert np.array_equal(alone, mixed)
    b.close()


def test_continuous_admission_gives_the_same_bits():
    pa, pb = problem(51, 16, 8), problem(52, 24, 11, crack=True)
    alone = {}
    for key, p in (("a", pa), ("b", pb)):
        b = E.Batch(E.SOLVE_MIXED)
        r = make(b, p)
        b.prepare([r])
        b.solve([r], tol_rel=1e-10, max_cycles=80)
        alone[key] = r.download(E.RVE.U)
        b.close()
    b = E.Batch(E.SOLVE_MIXED)
    ra, rb = make(b, pa), make(b, pb)
    b.prepare([ra, rb])
    b.submit(ra, 1e-10, 0.0, 80)
    for _ in range(3):
        b.step()
    b.submit(rb, 1e-10, 0.0, 80)  # joins while ra is in flight
    while b.step():
        pass
    assert ra.status()["converged"] and rb.status()["converged"]
    assert np.array_equal(ra.download(E.RVE.U), alone["a"])
    assert np.array_equal(rb.download(E.RVE.U), alone["b"])
    b.close()


def test_member_admission_respects_the_hbm_budget():
    need = E.Batch.member_bytes(32, E.SOLVE_MIXED, 16)
    ctx = E.Context(0, hbm_budget_bytes=int(need * 1.5))
    b = E.Batch(E.SOLVE_MIXED, ctx=ctx)
    r1 = b.rve(32, 16)
    with pytest.raises(E.ElaError) as ei:
        b.rve(32, 16)
    assert ei.value.code == 2
    r1.close()
    r2 = b.rve(32, 16)  # fits again once the first member is gone
    r2.close()
    b.close()
    ctx.close()


def test_members_of_any_side_solve_with_a_bicgstab_coarsest_level(oracle):
    """Sides whose coarsest level is beyond the dense inverse (36 -> 18 -> 9: 3000 unknowns; 27, odd: one
    level, solved by BiCGStab alone) solve in one batch with a dense-coarsest member (20 -> 10 -> 5) and
    match the oracle's MG with its BiCGStab bottom (the reference's bottom solver) to 1e-8."""
    probs = [problem(61, 36, 9), problem(62, 27, 6, crack=True), problem(63, 20, 5)]
    b = E.Batch(E.SOLVE_MIXED)
    rs = [make(b, p) for p in probs]
    b.prepare(rs)
    sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
    for r, p, st in zip(rs, probs, sts):
        n = p["n"]
        print("n=%d levels=%d cycles=%d res/res0=%.2e" % (n, r.levels(), st["iters"], st["res"] / st["res0"]))
        assert st["converged"] and st["res"] <= 1e-10 * st["res0"]
        bb = r.download(E.RVE.B)
        ref, st_or, _ = oracle.mg_solve(oracle.Grid((n, n, n)), oracle.MM, planes(r, p), all_disp(), bb,
                                        tol_rel=1e-11, coarse_min=2, bottom="bicgstab")
        assert st_or["res"] <= 1e-11 * st_or["res0"], st_or
        err = np.abs(r.download(E.RVE.U) - ref).max() / np.abs(ref).max()
        assert err <= 1e-8, (n, err)
    assert [r.levels() for r in rs] == [3, 1, 3]
    b.close()


def test_bicgstab_coarsest_member_is_the_same_alone_and_in_a_mix():
    target = problem(71, 44, 8, crack=True)  # 44 -> 22 -> 11: BiCGStab on 3 * 12^3 unknowns
    others = [problem(72, 27, 5), problem(73, 16, 7), problem(74, 36, 6)]
    b = E.Batch(E.SOLVE_MIXED)
    r = make(b, target)
    b.prepare([r])
    st_alone = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
    alone = r.download(E.RVE.U)
    b.close()
    b = E.Batch(E.SOLVE_MIXED)
    rs = [make(b, others[0]), make(b, target)] + [make(b, p) for p in others[1:]]
    b.prepare(rs)
    sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
    assert all(st["converged"] for st in sts)
    assert sts[1]["iters"] == st_alone["iters"]
    assert np.array_equal(rs[1].download(E.RVE.U), alone)
    b.close()


def test_serving_loop_results_match_alone_solves():
    """Requests served in the continuous-batching loop give bitwise the solution of the same
    request solved alone."""
    reqs = K.request_stream(rate=400.0, duration=0.08, seed=7, nmin=24, nmax=48)
    assert len(reqs) >= 10
    keep = (0, len(reqs) // 2, len(reqs) - 1)
    b = E.Batch(E.SOLVE_MIXED)
    res, sols = b.serve(reqs, tol_rel=1e-10, max_cycles=80, max_active=16, keep=keep)
    for i, r in enumerate(res):
        assert r["status"] == 0 and r["finished"] >= r["admitted"] >= 0.0, (i, r)
        assert r["res"] <= 1e-10 * r["res0"]
    for i in keep:
        q = reqs[i]
        r = make(b, q)
        b.prepare([r])
        st = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
        assert st["iters"] == res[i]["cycles"]
        assert np.array_equal(r.download(E.RVE.U), sols[i]), i
        r.close()
    b.close()


def mixed_member_bcs():
    """ZLO Displacement, ZHI Displacement in z and Traction in x, y; the x faces Traction and the y faces
    Neumann in every component; edges and corners follow their z face, else Traction."""
    t = np.zeros((26, 3), np.int32)
    for f, name in enumerate(O_FACES3):
        parts = name.split("_")
        if "ZLO" in parts:
            continue
        if "ZHI" in parts:
            t[f] = (1, 1, 0)
        elif parts == ["YLO"] or parts == ["YHI"]:
            t[f] = (3, 3, 3)
        else:
            t[f] = (1, 1, 1)
    return t


O_FACES3 = ["XLO", "YLO", "ZLO", "XHI", "YHI", "ZHI", "YLO_ZLO", "YLO_ZHI", "YHI_ZLO", "YHI_ZHI", "ZLO_XLO",
            "ZLO_XHI", "ZHI_XLO", "ZHI_XHI", "XLO_YLO", "XLO_YHI", "XHI_YLO", "XHI_YHI", "XLO_YLO_ZLO",
            "XLO_YLO_ZHI", "XLO_YHI_ZLO", "XLO_YHI_ZHI", "XHI_YLO_ZLO", "XHI_YLO_ZHI", "XHI_YHI_ZLO", "XHI_YHI_ZHI"]


@pytest.mark.parametrize("mode", [E.SOLVE_MIXED, E.SOLVE_FP64])
@pytest.mark.parametrize("bcs", ["displacement", "mixed"])
def test_marching_kernels_match_the_node_kernels(mode, bcs):
    """The LDS z-marching kernels of the finest levels and the one-thread-per-node kernels evaluate the
    same expressions; solves agree to rounding (member sizes with partial tiles in x, y and z), also with
    traction and Neumann faces, whose boundary rows divide by (FP64) or multiply by the inverse of (FP32)
    a diagonal that is not 1."""
    probs = [problem(61, 40, 12), problem(62, 24, 9, crack=True), problem(63, 32, 20)]
    sols, cyc = {}, {}
    for march in (0, 1):
        b = E.Batch(mode)
        b.set_option(E.Batch.MARCH, march)
        b.set_option(E.Batch.KZ, 16)
        rs = [make(b, p) for p in probs]
        if bcs == "mixed":
            for r, p in zip(rs, probs):
                r.SetBC(mixed_member_bcs())
                r.SetRhsAffine(p["G"])
        b.prepare(rs)
        sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
        sols[march] = [r.download(E.RVE.U) for r in rs]
        cyc[march] = [s["iters"] for s in sts]
        b.close()
    assert cyc[0] == cyc[1]
    for a, c in zip(sols[0], sols[1]):
        print("bitwise identical:", np.array_equal(a, c))
        assert np.abs(a - c).max() <= 1e-13 * np.abs(a).max()


def test_graph_replayed_batch_cycles_keep_the_bits():
    """OPT_GRAPH captures each step's F-cycle into a HIP graph and replays it while the active set stays
    the same; members retiring re-capture.  Two solves in a row (eager level-0 updates and residuals and a
    host synchronization between replays, captures and replays both exercised) give the stream launches'
    bits in both precisions; a batch with a BiCGStab coarsest member runs its cycles on the stream."""
    probs = [problem(81, 16, 8), problem(82, 24, 9, crack=True), problem(83, 32, 11), problem(84, 20, 5)]
    for mode in (E.SOLVE_MIXED, E.SOLVE_FP64):
        out = {}
        for graph in (0, 1):
            b = E.Batch(mode)
            b.set_option(E.OPT_GRAPH, graph)
            rs = [make(b, p) for p in probs]
            b.prepare(rs)
            sols = []
            for rep in range(2):
                for r in rs:
                    r.upload(E.RVE.U, np.zeros((3,) + (r.n + 1,) * 3))
                sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
                assert all(st["converged"] for st in sts)
                sols.append([r.download(E.RVE.U) for r in rs] + [[st["iters"] for st in sts]])
            out[graph] = sols
            if graph:
                caps, reps = b.graph_stats()
                assert caps >= 2 and reps > caps, (caps, reps)
            b.close()
        for a, c in zip(out[0], out[1]):
            for x, y in zip(a, c):
                assert np.array_equal(np.asarray(x), np.asarray(y))
    b = E.Batch(E.SOLVE_MIXED)
    b.set_option(E.OPT_GRAPH, 1)
    rs = [make(b, problem(85, 36, 6)), make(b, problem(86, 16, 5))]
    b.prepare(rs)
    assert all(st["converged"] for st in b.solve(rs, tol_rel=1e-10, max_cycles=80))
    assert b.graph_stats() == (0, 0)
    b.close()


def test_batch_autotune_keeps_the_bits():
    probs = [problem(71, 32, 10), problem(72, 24, 7, crack=True)]
    sols = []
    for tune in (False, True):
        b = E.Batch(E.SOLVE_MIXED)
        rs = [make(b, p) for p in probs]
        b.prepare(rs)
        if tune:
            times = b.autotune(rs, nrep=2)
            assert len(times) == 6 and all(t > 0 for t in times.values())
        b.solve(rs, tol_rel=1e-10, max_cycles=80)
        sols.append([r.download(E.RVE.U) for r in rs])
        b.close()
    for a, c in zip(*sols):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
def test_batched_gauss_jordan_one_launch_a_column_gives_the_launch_pair_bits(mode):
    """ELA_OPT_GJ on a batch: the members' coarsest inverses by Gauss-Jordan in one launch a column (the pivot
    search in every block over a compact candidate column, the row swaps as a permutation) or in pivot/eliminate
    launch pairs give the same bits, so the batched solves do.  Members whose dense coarsest levels differ in
    size (192, 648 and 1536 unknowns: 24, 40 and 28 cells a side), one with a crack factor."""
    probs = [problem(81, 24, 6), problem(82, 40, 9, crack=True), problem(83, 28, 7)]
    sols = []
    for gj in (0, 1):
        b = E.Batch(mode)
        b.set_option(E.OPT_GJ, gj)
        rs = [make(b, p) for p in probs]
        b.prepare(rs)
        b.solve(rs, tol_rel=1e-10, max_cycles=80)
        sols.append([r.download(E.RVE.U) for r in rs])
        b.close()
    for a, c in zip(*sols):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("mode", [E.SOLVE_FP64, E.SOLVE_MIXED])
def test_batched_per_grain_diagonal_keeps_the_bits(mode):
    """ELA_OPT_GRAIN_DIAG and ELA_OPT_PROLONG on a batch: the marching tiles of members without factors take an
    interior node's Jacobi diagonal from its grain (per member and level), or read the stored diagonal field;
    the prolongation runs one thread per coarse node (its 8 fine nodes) or one per fine node.  The solves agree
    bit for bit.  A member with a crack factor (which reads the field either way) rides along."""
    probs = [problem(84, 32, 6), problem(85, 40, 9, crack=True), problem(86, 48, 7)]
    sols = []
    for gd in (0, 1):
        b = E.Batch(mode)
        b.set_option(E.OPT_GRAIN_DIAG, gd)
        b.set_option(E.OPT_PROLONG, gd)
        rs = [make(b, p) for p in probs]
        b.prepare(rs)
        sts = b.solve(rs, tol_rel=1e-10, max_cycles=80)
        sols.append([(r.download(E.RVE.U), st["iters"]) for r, st in zip(rs, sts)])
        b.close()
    for (a, ca), (c, cc) in zip(*sols):
        assert ca == cc
        assert np.array_equal(a, c)


def test_mixed_batch_smooths_traction_boundaries_like_fp64():
    """The FP32 levels hold the inverse Jacobi diagonal: the marching launches' shell rows must multiply by it
    (a division there scaled the update of every traction boundary node by D^2).  With traction faces the
    mixed batched solve then needs the FP64 solve's cycles, to the same solution."""
    out = {}
    for mode in (E.SOLVE_FP64, E.SOLVE_MIXED):
        b = E.Batch(mode)
        p = problem(51, 16, 8)
        r = make(b, p)
        bct = all_disp()
        bct[[0, 1, 3, 4]] = K.TRAC
        bct[6:, 0] = K.TRAC
        r.SetBC(bct)
        b.prepare([r])
        st = b.solve([r], tol_rel=1e-10, max_cycles=100)[0]
        assert st["converged"]
        out[mode] = (r.download(E.RVE.U), st["iters"])
        b.close()
    (u64, i64), (umx, imx) = out[E.SOLVE_FP64], out[E.SOLVE_MIXED]
    assert imx <= i64 + 2, (imx, i64)
    assert np.abs(umx - u64).max() <= 1e-8 * np.abs(u64).max()


def failing_mix():
    """A small request stream with two failing requests: one with a non-finite stiffness (NaN C11) and
    one with a cycle limit of 1."""
    reqs = K.request_stream(rate=400.0, duration=0.05, seed=11, nmin=24, nmax=40)
    assert len(reqs) >= 8
    reqs[2] = dict(reqs[2], C11=float("nan"))
    reqs[5] = dict(reqs[5], max_cycles=1)
    return reqs


def alone(b, q):
    r = make(b, q)
    b.prepare([r])
    st = b.solve([r], tol_rel=1e-10, max_cycles=80)[0]
    u = r.download(E.RVE.U)
    r.close()
    return st, u


def test_serving_loop_isolates_failing_requests():
    """ela_serve: the NaN request ends with ELA_ERR_NAN (its set-up fails alone: the admitted group is set
    up member by member), the one with a cycle limit of 1 with ELA_ERR_NONCONVERGED; every other request
    converges with the cycles and bits it has alone."""
    reqs = failing_mix()
    keep = tuple(i for i in range(len(reqs)) if i not in (2, 5))
    b = E.Batch(E.SOLVE_MIXED)
    res, sols = b.serve(reqs, tol_rel=1e-10, max_cycles=80, max_active=16, keep=keep)
    assert res[2]["status"] == 5, res[2]
    assert res[5]["status"] == 4 and res[5]["cycles"] == 1, res[5]
    for i in keep:
        assert res[i]["status"] == 0, (i, res[i])
        st, u = alone(b, reqs[i])
        assert st["iters"] == res[i]["cycles"]
        assert np.array_equal(u, sols[i]), i
    b.close()


def test_online_server_admits_requests_while_others_run():
    """ela_server_*: requests submitted in three waves while earlier ones are in flight are admitted into
    the running loop; the failing ones get their own status and the others the bits they have alone."""
    import time
    reqs = failing_mix()
    b = E.Batch(E.SOLVE_MIXED)
    sv = E.Server(b, tol_rel=1e-10, max_cycles=80, max_active=16)
    sols, tickets = {}, {}
    waves = [range(0, 3), range(3, 6), range(6, len(reqs))]
    for w in waves:
        for i in w:
            sols[i] = np.zeros((3,) + (reqs[i]["n"] + 1,) * 3)
            tickets[i] = sv.submit(reqs[i], u_out=sols[i])
        time.sleep(0.02)
    res, t0 = {}, time.time()
    while len(res) < len(reqs) and time.time() - t0 < 60:
        for i, t in tickets.items():
            if i not in res:
                r = sv.poll(t)
                if r is not None:
                    res[i] = r
        time.sleep(0.005)
    sv.stop()
    assert len(res) == len(reqs)
    assert res[2]["status"] == 5 and res[5]["status"] == 4
    overlap = False
    for i in range(len(reqs)):
        if i in (2, 5):
            continue
        assert res[i]["status"] == 0, (i, res[i])
        st, u = alone(b, reqs[i])
        assert st["iters"] == res[i]["cycles"]
        assert np.array_equal(u, sols[i]), i
    # some request was admitted while another was still in flight (continuous batching)
    fin = [(res[i]["admitted"], res[i]["finished"]) for i in res if res[i]["status"] in (0, 4)]
    overlap = any(a2 > a1 and a2 < f1 for a1, f1 in fin for a2, _ in fin)
    assert overlap
    b.close()
