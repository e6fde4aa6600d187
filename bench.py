"""Benchmark of the nodal elasticity solve on one MI355X.

Workload (default): 3D single level, 256^3 cells, cubic (C11, C12, C44) =
(1.68, 1.21, 0.75) Voronoi polycrystal of 64 grains with Haar-random
orientations (fixed seed), uniaxial strain 0.01 along z.  A step is one full
solve from u = 0 to a relative residual of 1e-10: set-up of the operator
(stiffness coarsening, diagonals, coarsest inverse) plus the multigrid
iterations, with b and u resident in HBM.  The headline value is the
mixed-precision (FP32 smoothing / FP64 refinement) solves/s; the plain FP64
GPU solver is measured in the same run.

Prints one JSON line.  Usage: python bench.py [--steps K] [--warmup W] [--n 256]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_mode(op, b, u, mode, steps, warmup, tol, timing=False):
    import anisotropicfracture_amd as ela
    stats = None
    for _ in range(warmup):
        u.zero()
        op.prepareForSolve()
        rc, stats = op.solve_device(b, u, tol, 0.0, mode, 400)
    op.set_timing(timing)
    t0 = time.perf_counter()
    per = []
    for _ in range(steps):
        ts = time.perf_counter()
        u.zero()
        op.prepareForSolve()
        rc, stats = op.solve_device(b, u, tol, 0.0, mode, 400)
        per.append(time.perf_counter() - ts)
        if rc not in (0, 4):
            raise RuntimeError("solve failed: rc=%d %s" % (rc, ela.lib().ela_last_error()))
    t = time.perf_counter() - t0
    times = op.kernel_times() if timing else None
    op.set_timing(False)
    return t, per, stats, rc, times


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity}


def cpu_and_checks(n_full, n_small, plain_cycles, cases):
    """The oracle leg of the bench: the CPU baseline and the per-config error of the GPU results.

    CPU baseline: the oracle (the FP64 C++ restatement of the reference algorithm: plain MG F-cycles,
    Jacobi smoothing, OpenMP on this host's cores) solves the polycrystal at n_small^3 to 1e-10 in full;
    at the full n_full^3 it runs its set-up and 2 F-cycles, and its solve time there is set-up +
    plain_cycles x (time per F-cycle), plain_cycles being the FP64 GPU solver's cycle count for the same
    algorithm (per-node planes, plain MG) at n_full^3.  One F-cycle at n_small^3 on one thread gives the
    single-thread rate.

    Checks: the GPU solutions of `cases` against the oracle (the checker only; nothing here is timed
    for the GPU)."""
    import scipy.sparse.linalg as spl
    from anisotropicfracture_amd import configs as K
    from oracle import oracle as O
    from oracle import amr as A
    O.build()
    threads = O.get_threads()
    out = {"cpu": cpu_info(), "threads": threads}
    errs = {}

    def polyC(cfg):
        gid = O.voronoi_ids(cfg["seeds"], cfg["n"], K.VORONOI_Q)
        tab = np.stack([O.cubic_define(cfg["C11"], cfg["C12"], cfg["C44"], R) for R in cfg["R"]])
        C = np.empty((21,) + gid.shape)
        for q in range(21):
            C[q] = tab[:, q][gid]
        return C

    # full solve at n_small^3 (all threads), then its reference solution to 1e-13 and one single-thread cycle
    cfg = K.polycrystal(n_small)
    C = polyC(cfg)
    g = O.Grid(cfg["n"])
    b = O.bc_init(g, cfg["bc_type"], cfg["bc_value"], g.zeros())
    t0 = time.perf_counter()
    u_s, st_s, _ = O.mg_solve(g, O.MM, C, cfg["bc_type"], b, tol_rel=1e-10, max_iter=400)
    t_small = time.perf_counter() - t0
    u_ref, st_ref, _ = O.mg_solve(g, O.MM, C, cfg["bc_type"], b, u0=u_s, tol_rel=1e-12 * st_s["res0"] / st_s["res"],
                                  max_iter=60)
    O.set_threads(1)
    _, st1, _ = O.mg_solve(g, O.MM, C, cfg["bc_type"], b, tol_rel=1e-300, max_iter=1)
    O.set_threads(threads)
    out["small"] = {"n": n_small, "seconds": t_small, "cycles": st_s["iters"], "converged": st_s["converged"],
                    "setup_s": st_s["setup_s"], "single_thread_cycle_s": st1["iter_s"],
                    "single_thread_setup_s": st1["setup_s"]}
    if "poly_small" in cases:
        u_gpu = cases["poly_small"]
        errs["cfg2_64"] = {"n": n_small, "rel_disp_err": float(np.abs(u_gpu - u_ref).max() / np.abs(u_ref).max()),
                           "oracle_rel_residual": float(np.abs(b - O.fapply(g, O.MM, C, cfg["bc_type"], u_gpu)).max()
                                                        / np.abs(b).max()),
                           "reference": "oracle MG to %.1e" % (st_ref["res"] / st_s["res0"])}
    del C, u_s, u_ref
    # the full size: set-up and 2 F-cycles, and the oracle residual of the GPU headline solution
    cfg = K.polycrystal(n_full)
    C = polyC(cfg)
    g = O.Grid(cfg["n"])
    b = O.bc_init(g, cfg["bc_type"], cfg["bc_value"], g.zeros())
    _, st2, _ = O.mg_solve(g, O.MM, C, cfg["bc_type"], b, tol_rel=1e-300, max_iter=2)
    t_cycle = st2["iter_s"] / max(1, st2["iters"])
    t_solve = st2["setup_s"] + plain_cycles * t_cycle
    out["full"] = {"n": n_full, "setup_s": st2["setup_s"], "cycle_s": t_cycle, "cycles_timed": st2["iters"],
                   "cycles_needed": plain_cycles, "solve_s": t_solve}
    if "poly_full" in cases:
        u_gpu = cases["poly_full"]
        r = b - O.fapply(g, O.MM, C, cfg["bc_type"], u_gpu)
        errs["cfg2"] = {"n": n_full, "oracle_rel_residual": float(np.abs(r).max() / np.abs(b).max()),
                        "note": "the oracle's residual of the GPU mixed solution (same inputs, FP64)"}
        del r
    del C
    out["value"] = 1.0 / t_solve
    # config 1: 2D inclusion, sparse direct solve of the oracle-assembled operator
    if "inclusion" in cases:
        nn, u_gpu = cases["inclusion"]
        (nx, ny), C1, t1, v1 = K.inclusion_2d(nn)
        nest = A.Nest((nx, ny), [])
        b1 = nest.bc_init(0, t1, v1, np.zeros(nest.shape(0, 2)))
        M = nest._level_matrix(0, O.ISO, C1, t1, False).tocsc()
        lu = spl.splu(M)
        g1 = O.Grid((nx, ny))
        u_lu = lu.solve(b1.ravel()).reshape(b1.shape)
        res_lu = np.abs(b1 - O.fapply(g1, O.ISO, C1, t1, u_lu)).max() / np.abs(b1).max()
        # one step of iterative refinement in FP64: the factorization alone leaves a residual of 2e-8 and an
        # error of the same size; refined, the solution is fixed to round-off (a second step moves it by 1e-15)
        u_dir = u_lu + lu.solve((b1 - O.fapply(g1, O.ISO, C1, t1, u_lu)).ravel()).reshape(b1.shape)
        res_dir = np.abs(b1 - O.fapply(g1, O.ISO, C1, t1, u_dir)).max() / np.abs(b1).max()
        e = {}
        for key, u in u_gpu.items():
            e[key] = {"rel_disp_err": float(np.abs(u - u_dir).max() / np.abs(u_dir).max()),
                      "oracle_rel_residual": float(np.abs(b1 - O.fapply(g1, O.ISO, C1, t1, u)).max() / np.abs(b1).max())}
        e["n"] = nn
        e["direct_solve_rel_residual"] = float(res_lu)
        e["refined_direct_solve_rel_residual"] = float(res_dir)
        e["reference"] = "sparse LU of the oracle-assembled operator + one FP64 refinement step"
        e["note"] = ("FP64 floor: the LU solve reaches %.1e and the refined one %.1e, so the GPU solves run to "
                     "3e-9" % (res_lu, res_dir))
        errs["cfg1"] = e
    # config 3: the oracle's composite residual (solutionResidual + reflux, masked) of the GPU solution
    if "amr" in cases:
        a_cfg, Cs, bs, us = cases["amr"]
        nest = A.Nest(a_cfg["n"], a_cfg["boxes"])
        r, _ = nest.residual(O.ISO, Cs, a_cfg["bc_type"], bs, us)
        errs["cfg3"] = {"base": a_cfg["n"][0], "oracle_rel_residual": float(max(np.abs(x).max() for x in r) /
                                                                           max(np.abs(x).max() for x in bs)),
                        "note": "the oracle's composite residual of the GPU mixed solution; the composite "
                                "displacement is compared with a sparse direct solve in tests/test_gpu_amr_composite.py"}
    # config 4: a sample of members against the oracle MG (coarsening to 2 cells, as the members)
    if "rves" in cases:
        worst_e = worst_r = 0.0
        for m, u in cases["rves"]:
            n = m["n"]
            gr = O.Grid((n, n, n))
            c4 = {"n": (n, n, n), "seeds": m["seeds"], "R": m["R"], "C11": K.C11, "C12": K.C12, "C44": K.C44}
            C4 = polyC(c4)
            bct = np.zeros((26, 3), np.int32)
            x = gr.coords()
            bh = np.zeros(gr.shape(3))
            bnd = np.zeros(x[0].shape, bool)
            bnd[[0, -1], :, :] = bnd[:, [0, -1], :] = bnd[:, :, [0, -1]] = True
            for p in range(3):
                bh[p][bnd] = sum(m["G"][p, d] * x[d][bnd] for d in range(3))
            u_or, st4, _ = O.mg_solve(gr, O.MM, C4, bct, bh, tol_rel=1e-12, max_iter=80, coarse_min=2)
            worst_e = max(worst_e, float(np.abs(u - u_or).max() / np.abs(u_or).max()))
            worst_r = max(worst_r, float(np.abs(bh - O.fapply(gr, O.MM, C4, bct, u)).max() / np.abs(bh).max()))
        errs["cfg4_sample"] = {"members": [m["n"] for m, _ in cases["rves"]], "max_rel_disp_err": worst_e,
                               "max_oracle_rel_residual": worst_r, "reference": "oracle MG to 1e-12"}
    out["per_config_error"] = errs
    return out


def batched_rves(count, sample, tol, graphs=1):
    """Batched RVEs: `count` independent RVEs (sides 32/48/64/96, 8-32 grains, fixed-corner uniaxial z
    strain 0.01).  Batched mixed and batched FP64 solve every RVE, one launch per cycle kernel over all
    of them (mixed also with each step's F-cycle replayed from a captured HIP graph); the plain FP64 GPU
    solver (ela_solve on one hierarchy per size, per-node stiffness planes, one RVE at a time: the batch's
    MG iteration, same coarse operators) solves `sample` RVEs, the same number of each side, and its
    rate for the whole set is the per-side mean time weighted by the set's side counts.  Timed per RVE:
    stiffness evaluation on the device, set-up (coarse stiffness, diagonals, coarsest inverse) and the
    solve."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    cfg = K.rve_batch(count)
    out = {"rves": count, "nodes": int(sum((m["n"] + 1) ** 3 for m in cfg))}
    cycles = {}

    def load(b, rs, ms):
        for r, m in zip(rs, ms):
            r.SetModelVoronoiCubic(m["seeds"], K.VORONOI_Q, m["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
            r.SetRhsAffine(m["G"])

    variants = [(ela.SOLVE_MIXED, "mixed", 0), (ela.SOLVE_MIXED, "mixed_graph", 1), (ela.SOLVE_FP64, "fp64", 0)]
    for mode, key, graph in [v for v in variants if graphs or not v[2]]:
        b = ela.Batch(mode)
        b.set_option(ela.OPT_GRAPH, graph)
        # untimed warm-up batch: first launches of every kernel, and the marching tile shape picked on
        # a launch large enough to fill the GPU
        nw = min(64, count)
        warm = [b.rve(m["n"], max_grains=32) for m in cfg[:nw]]
        load(b, warm, cfg[:nw])
        b.prepare(warm)
        tuned = {"%d/%d" % k: v for k, v in b.autotune(warm, nrep=3).items()} if not graph else None
        b.solve(warm, tol_rel=tol, max_cycles=100)
        for r in warm:
            r.close()
        rs = [b.rve(m["n"], max_grains=32) for m in cfg]
        b.set_timing(key == "mixed")
        t0 = time.perf_counter()
        load(b, rs, cfg)
        b.prepare(rs)
        sts = b.solve(rs, tol_rel=tol, max_cycles=100, raise_on_fail=False)
        t = time.perf_counter() - t0
        it = np.array([q["iters"] for q in sts])
        cycles[key] = it
        ent = {"solves_per_s": count / t, "seconds": t, "cycles_mean": float(it.mean()), "cycles_max": int(it.max()),
               "converged": int(sum(q["converged"] for q in sts)),
               "max_rel_residual": float(max(q["res"] / q["res0"] for q in sts)), "device_GB": b.bytes() / 1e9,
               "autotune_ms_per_sweep": tuned}
        if graph:
            ent["graph_captures_replays"] = b.graph_stats()
        if key == "mixed":
            sm = b.smoother_times()
            ent["smoother"] = {"ms": sm["ms"], "launches": sm["launches"], "node_sweeps": sm["nodes"],
                               "Gnodes_per_s": sm["nodes"] / (sm["ms"] * 1e-3) / 1e9}
        out["batched_" + key] = ent
        log("bench: batched %s %.1f solves/s (%.2f s, cycles %.1f)" % (key, count / t, t, it.mean()))
        b.close()
    # plain FP64 GPU solver, one RVE at a time (hierarchies reused per size), the same number of each side
    sides = sorted(set(m["n"] for m in cfg))
    per = max(1, sample // len(sides))
    pick = [i for s_ in sides for i in [j for j, m in enumerate(cfg) if m["n"] == s_][:per]]
    ops, tot, its, tside = {}, 0.0, [], {s_: [] for s_ in sides}
    for i in pick:
        m = cfg[i]
        n = m["n"]
        if n not in ops:
            op = ela.Elastic((n, n, n), ela.MM, mode=ela.SOLVE_FP64, coarse_min=2)
            op.SetBC(np.zeros((26, 3), np.int32), np.zeros((26, 3)))
            ops[n] = (op, op.field(0, ela.F64, 3), op.field(0, ela.F64, 3))
        op, bf, uf = ops[n]
        x = np.arange(n + 1) / n
        z = np.broadcast_to(x[:, None, None], (n + 1,) * 3)
        bh = np.zeros((3,) + (n + 1,) * 3)
        bnd = np.ones((n + 1,) * 3, bool)
        bnd[1:-1, 1:-1, 1:-1] = False
        bh[2][bnd] = m["G"][2, 2] * z[bnd]
        bf.upload(bh)
        uf.zero()
        t0 = time.perf_counter()
        op.SetModelVoronoiCubic(m["seeds"], K.VORONOI_Q, m["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
        op.SetStiffnessMode(1)
        op.prepareForSolve()
        rc, st = op.solve_device(bf, uf, tol, 0.0, ela.SOLVE_FP64, 200)
        dt = time.perf_counter() - t0
        tot += dt
        tside[n].append(dt)
        its.append(st["iters"])
    for op, _, _ in ops.values():
        op.close()
    est = sum(sum(1 for m in cfg if m["n"] == s_) * float(np.mean(tside[s_])) for s_ in sides)
    same = [a == cycles["fp64"][i] for a, i in zip(its, pick)]
    out["plain_fp64"] = {"solves_per_s": count / est, "sample": len(pick), "per_side": per,
                         "s_per_solve_by_side": {str(s_): float(np.mean(tside[s_])) for s_ in sides},
                         "estimate": "sum over sides of (RVEs of that side in the set) x (mean sampled time)",
                         "cycles_mean": float(np.mean(its)), "stiffness": "grain form (grain ids and stiffness injected onto the coarse levels, as the batch)",
                         "cycles_same_as_batched_fp64": bool(all(same)),
                         "cycles_equal_fraction": float(np.mean(same))}
    out["batched_mixed_over_plain_fp64"] = out["batched_mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    out["batched_fp64_over_plain_fp64"] = out["batched_fp64"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    log("bench: plain fp64 %.1f solves/s (sample %d); batched mixed / plain fp64 = %.2f"
        % (out["plain_fp64"]["solves_per_s"], len(pick), out["batched_mixed_over_plain_fp64"]))
    return out


def amr_crack(n, steps, tol, krylov, keep=None):
    """The mode-I crack on a 3-level AMR nest (128^3 base) (composite solve), mixed and FP64.

    A step is prepareForSolve + one composite solve from u = 0 with b and the model resident.  With
    keep (a dict), the inputs and the mixed solution per level go into keep["amr"] for the checker."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    cfg = K.crack_nest(n)
    shapes = sorted({tuple(h[d] - l[d] for d in range(3)) for lv in cfg["boxes"] for l, h in lv})
    out = {"base": n, "regions": [[list(lo), list(hi)] for lo, hi in cfg["regions"]],
           "max_grid_size": cfg["max_grid_size"],
           "boxes_per_level": [int(np.prod([-(-n // cfg["max_grid_size"])] * 3))] + [len(b) for b in cfg["boxes"]],
           "refined_box_shapes": [list(t) for t in shapes], "g": list(cfg["g"]), "width": cfg["width"]}
    sols = {}
    for key, mode in (("fp64", ela.SOLVE_FP64), ("mixed", ela.SOLVE_MIXED)):
        a = ela.ElasticAMR(cfg["n"], cfg["boxes"], ela.ISO, mode=mode, coarse_min=4)
        a.SetBC(cfg["bc_type"], cfg["bc_value"])
        a.SetKrylov(krylov)
        nodes, Cs, bs = [], [], []
        for l in range(a.nlev):
            X = a.coords(l)
            Cs.append(K.crack_nest_model(cfg, X[0], X[1], X[2]))
            bs.append(a.bc_init(l))
            a.SetModel(l, Cs[l])
            a.upload(a.B, l, bs[l])
            nodes.append(int(np.prod(a.shape(l)[1:])))
        zero = [np.zeros(a.shape(l)) for l in range(a.nlev)]
        times, st = [], None
        for it in range(steps + 1):  # the first solve is untimed (first launches)
            for l in range(a.nlev):
                a.upload(a.U, l, zero[l])
            t0 = time.perf_counter()
            a.prepareForSolve()
            st = a.solve_resident(tol_rel=tol, max_iter=400, raise_on_fail=False)
            t1 = time.perf_counter()
            if it:
                times.append(t1 - t0)
        sols[key] = [a.download(a.U, l) for l in range(a.nlev)]
        if keep is not None and key == "mixed":
            keep["amr"] = (cfg, Cs, bs, sols[key])
        out[key] = {"solves_per_s": len(times) / sum(times), "s_per_solve": sum(times) / len(times),
                    "iters": st["iters"], "rel_residual": st["res"] / st["res0"], "converged": bool(st["converged"]),
                    "device_GB": a.bytes() / 1e9}
        out["nodes_per_level"] = nodes
        log("bench: amr crack %s %.3f s/solve iters=%d rel=%.2e" % (key, sum(times) / len(times), st["iters"],
                                                                     st["res"] / st["res0"]))
        a.close()
    out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    scale = max(np.abs(x).max() for x in sols["fp64"])
    out["rel_disp_diff_mixed_vs_fp64"] = max(np.abs(x - y).max() for x, y in zip(sols["mixed"], sols["fp64"])) / scale
    return out


def inclusion_2d(n, steps, keep):
    """Config 1: the 2D inclusion at n^2 (isotropic E = 210, nu = 0.3, r = 0.2 inclusion of 10x stiffness),
    FGMRES(24) + MG, FP64 and mixed; a step is prepareForSolve + one solve.  The tolerance is 3e-9: this
    operator's FP64 residual floor (a sparse direct solve of it stops at 2e-8 relative, the FGMRES
    iterates at about 2.2e-9), far above 1e-10."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    (nx, ny), C, t, v = K.inclusion_2d(n)
    op = ela.Elastic((nx, ny), ela.ISO, mode=ela.ALLOC_BOTH)
    op.SetModel(C)
    op.SetBC(t, v)
    op.SetKrylov(24)
    op.prepareForSolve()
    bh = op.bc_init()
    b = op.field(0, ela.F64).upload(bh)
    u = op.field(0, ela.F64)
    tol = 3e-9
    out, sols = {"n": n, "tol_rel": tol}, {}
    for key, mode in (("fp64", ela.SOLVE_FP64), ("mixed", ela.SOLVE_MIXED)):
        t_, per, st, rc, _ = run_mode(op, b, u, mode, steps, 1, tol)
        sols[key] = u.download()
        out[key] = {"solves_per_s": steps / t_, "iters": st["iters"], "rel_residual": st["res"] / st["res0"],
                    "converged": bool(st["converged"])}
    op.close()
    keep["inclusion"] = (n, sols)
    log("bench: inclusion 2D %d^2 fp64 %.1f solves/s, mixed %.1f solves/s" % (n, out["fp64"]["solves_per_s"],
                                                                          out["mixed"]["solves_per_s"]))
    return out


def small_solutions(n_small, nrves, tol, keep):
    """GPU solutions for the checker: the polycrystal at n_small^3 (mixed, grain form, FGMRES(6)) and the
    first members of the RVE batch of sides 32 and 48 (batched mixed)."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    cfg = K.polycrystal(n_small)
    op = ela.Elastic(cfg["n"], ela.MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    op.SetBC(cfg["bc_type"], cfg["bc_value"])
    op.SetKrylov(6)
    op.SetStiffnessMode(1)
    op.prepareForSolve()
    bh = op.bc_init()
    u, st = ela.Linear(op).solve(np.zeros_like(bh), bh, tol_rel=tol, tol_abs=0.0, mode=ela.SOLVE_MIXED, max_iter=400)
    keep["poly_small"] = u
    op.close()
    ms = [m for m in K.rve_batch(64) if m["n"] in (32, 48)][:nrves]
    b = ela.Batch(ela.SOLVE_MIXED)
    rs = [b.rve(m["n"], max_grains=32) for m in ms]
    for r, m in zip(rs, ms):
        r.SetModelVoronoiCubic(m["seeds"], K.VORONOI_Q, m["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
        r.SetRhsAffine(m["G"])
    b.prepare(rs)
    b.solve(rs, tol_rel=tol, max_cycles=100)
    keep["rves"] = [(m, r.download(ela.RVE.U)) for r, m in zip(rs, ms)]
    for r in rs:
        r.close()
    b.close()


def serving(seconds, max_wall, tol, budget_gb=200.0, max_active=256, rate=200.0, check=True):
    """The serving stream: Poisson(rate) requests (200/s), arrivals in the first `seconds` s, served by
    the continuous-batching loop (ela_serve) under a `budget_gb` HBM budget, stopped after `max_wall` s.
    With check, two requests are re-solved alone afterwards and compared bit for bit."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    reqs = [q for q in K.request_stream(rate=rate) if q["arrival"] < seconds]
    ctx = ela.Context(0, int(budget_gb * 1e9))
    b = ela.Batch(ela.SOLVE_MIXED, ctx=ctx)
    keep = (0, min(40, len(reqs) - 1)) if check else ()
    res, sols = b.serve(reqs, tol_rel=tol, max_cycles=100, max_active=max_active, max_seconds=max_wall, keep=keep)
    summary = dict(b.last_serve)
    done = [(r, q) for r, q in zip(res, reqs) if r["status"] == 0 and r["finished"] >= 0]
    lat = np.array([r["finished"] - q["arrival"] for r, q in done])
    wall = max(r["finished"] for r, _ in done) if done else float("nan")
    same = True if check else None
    for i in keep:
        if res[i]["status"] != 0:
            same = None
            continue
        q = reqs[i]
        r = b.rve(q["n"], max_grains=len(q["seeds"]), with_factor=q["crack"] is not None)
        r.SetModelVoronoiCubic(q["seeds"], K.VORONOI_Q, q["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
        if q["crack"] is not None:
            p, nrm, tv, w, g = q["crack"]
            r.SetCrack(p, nrm, tv, w, *g)
        r.SetRhsAffine(q["G"])
        b.prepare([r])
        b.solve([r], tol_rel=tol, max_cycles=100)
        same = same and bool(np.array_equal(r.download(ela.RVE.U), sols[i]))
        r.close()
    b.close()
    ctx.close()
    nodes = np.array([(q["n"] + 1) ** 3 for _, q in done], dtype=np.float64)
    out = {"offered_rate_per_s": rate, "arrival_window_s": seconds, "requests": len(reqs), "served": len(done),
           "wall_s": wall, "solves_per_s": len(done) / wall if done else 0.0,
           "p50_ms": float(np.percentile(lat, 50) * 1e3) if done else None,
           "p99_ms": float(np.percentile(lat, 99) * 1e3) if done else None,
           "served_Mnodes_per_s": float(nodes.sum() / wall / 1e6) if done else None,
           "budget_GB": budget_gb, "max_active": max_active, "alone_vs_mix_bitwise": same,
           "sides": "requested side rounded to the nearest multiple of 4 (the reference's amr.blocking_factor)",
           "max_side_deviation": float(max(abs(q["n"] - q["n_requested"]) / q["n_requested"] for q in reqs)),
           "bicgstab_bottom_requests": int(sum(1 for q in reqs if K.dense_coarsest_unknowns(q["n"]) > 1536)),
           "not_served_by_limit": int(sum(1 for r in res if r["status"] == -1)),
           "time_breakdown_s": summary}
    log("bench: serving at %.0f/s: %d/%d requests in %.1f s: %.1f solves/s p50 %.0f ms p99 %.0f ms"
        % (rate, len(done), len(reqs), wall, out["solves_per_s"], out["p50_ms"] or -1, out["p99_ms"] or -1))
    return out


def pmc_traffic(fetch_csv, write_csv, n, tile, out_path):
    """HBM bytes per launch of the level-0 kernels from two `rocprofv3 --pmc` runs of this bench (one with
    FETCH_SIZE, one with WRITE_SIZE; their counter_collection.csv files), each calibrated by what the same
    counter reads for the copy probe of known size (1 GiB read + 1 GiB written a launch) in the same run.
    Writes the per-kernel table that the bench line's roofline.traffic reads."""
    import csv
    import re
    nodes = (n + 1) ** 3
    probe = float(1 << 30)
    # (pattern on the kernel name, description, algorithmic bytes per node); level 0 only (grid size filter).
    # The Jacobi sweep reads x, b and the grain id and writes x; the diagonal of an interior node comes from
    # its grain (LDS), the shell's nodes read theirs (12 B more on (n+1)^3 - (n-1)^3 nodes).
    jac_node = 4 * 9 + 2 + 12.0 * ((n + 1) ** 3 - (n - 1) ** 3) / nodes
    kinds = [(r"k_grain_sweep<float, 0,", "k_grain_sweep<float,0> (level-0 Jacobi sweep, grain form: shell + marching tiles)", jac_node),
             (r"k_grain_sweep<double, 3,", "k_grain_sweep<double,3> (level-0 Fapply FP64, grain form)", 8 * 6 + 2),
             (r"k_grain_sweep<float, 3,", "k_grain_sweep<float,3> (level-0 Fapply FP32, grain form)", 4 * 6 + 2),
             (r"k_fapply<double, 1, 3, ela::(Masked)?PlaneAcc", "k_fapply<double,MM,3,planes> (level-0 Fapply FP64, per-node planes)", 8 * 27),
             (r"k_fapply<float, 1, 3, ela::(Masked)?PlaneAcc", "k_fapply<float,MM,3,planes> (level-0 Fapply FP32, per-node planes)", 4 * 27)]

    def rows(path, counter):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") == counter:
                    yield r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"]) * 1024.0

    def per_kernel(path, counter):
        copy, acc = [], {}
        for name, grid, v in rows(path, counter):
            if re.search(r"\bk_copy\b|k_copy<|k_copy\(", name):
                copy.append(v)
            for pat, desc, _ in kinds:
                if re.search(pat, name):
                    acc.setdefault(desc, {}).setdefault(grid, []).append(v)
        return copy, acc

    fc, fa = per_kernel(fetch_csv, "FETCH_SIZE")
    wc, wa = per_kernel(write_csv, "WRITE_SIZE")
    if not fc or not wc:
        raise SystemExit("no copy-probe dispatches in the counter files: run the bench with --only poly")
    ff, wf = probe / float(np.median(fc)), probe / float(np.median(wc))
    table = {}
    for pat, desc, per_node in kinds:
        if desc not in fa or desc not in wa:
            continue
        grid = max(fa[desc])  # the level-0 launches have the largest grid of their kernel
        fv, wv = fa[desc][grid], wa[desc].get(grid, [])
        if not wv:
            continue
        fb, wb = float(np.mean(fv)) * ff, float(np.mean(wv)) * wf
        table[desc] = {"grid_threads": grid, "fetch_bytes": fb, "write_bytes": wb, "algorithmic_bytes": per_node * nodes,
                       "traffic_over_algorithmic": (fb + wb) / (per_node * nodes), "dispatches": len(fv)}
    out = {"n": n, "nodes": nodes, "tile": tile,
           "source": "rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of: python3 bench.py --only poly "
                     "--no-cpu --graphs 0 --steps 1 --warmup 1 --n %d --tile %s" % (n, tile),
           "calibration": {"kernel": "k_copy (ela_probe_copy variant 0, 1 GiB read + 1 GiB written per launch)",
                           "fetch_factor": ff, "write_factor": wf},
           "kernels": table}
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


def trace_summary(trace_csv, top=40, by_grid=True, segment=None, gap_ms=0.5):
    """Markdown table of a `rocprofv3 --kernel-trace` run's kernel_trace.csv grouped by kernel and grid size
    (a level's launches are told apart by their grid), or by kernel only: calls, total ms, average us, share.
    segment K: only the K-th run of dispatches separated from the next by more than gap_ms of idle GPU (a
    solve's launches follow each other closely; the host work between two solves leaves a gap), with its
    span and the idle time between its kernels."""
    import csv
    with open(trace_csv) as fh:
        rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
    if segment is not None:
        segs = [[rows[0]]]
        for a, b in zip(rows, rows[1:]):
            if (int(b["Start_Timestamp"]) - int(a["End_Timestamp"])) * 1e-6 > gap_ms:
                segs.append([])
            segs[-1].append(b)
        for i, sg in enumerate(segs):
            busy = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in sg) * 1e-6
            span = (int(sg[-1]["End_Timestamp"]) - int(sg[0]["Start_Timestamp"])) * 1e-6
            if len(sg) > 100:
                print("segment %d: %d dispatches, span %.2f ms, kernel time %.2f ms" % (i, len(sg), span, busy))
        rows = segs[segment]
        span = (int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])) * 1e-6
        print("\nSegment %d: %d dispatches over %.2f ms.\n" % (segment, len(rows), span))
    acc = {}
    for r in rows:
        name = r["Kernel_Name"].replace("void ", "").replace("ela::(anonymous namespace)::", "").replace("ela::", "")
        if not name.startswith("__amd"):
            name = name.split("(")[0]
        grid = (r.get("Grid_Size_X") or r.get("Grid_Size") or "?") if by_grid else "all"
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        e = acc.setdefault((name, grid), [0, 0.0])
        e[0] += 1
        e[1] += d
    total = sum(v[1] for v in acc.values())
    print("Total kernel time %.1f ms over %d kernel/grid pairs.\n" % (total * 1e-3, len(acc)))
    print("| kernel | grid threads | calls | total ms | avg us | % |")
    print("|---|---|---|---|---|---|")
    for (name, grid), (n, t) in sorted(acc.items(), key=lambda kv: -kv[1][1])[:top]:
        print("| `%s` | %s | %d | %.2f | %.1f | %.1f |" % (name, grid, n, t * 1e-3, t / n, 100.0 * t / total))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--trace-summary":
        # python bench.py --trace-summary KERNEL_TRACE.csv [TOP [name|grid [SEGMENT [GAP_MS]]]]
        trace_summary(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 40,
                      by_grid=not (len(sys.argv) > 4 and sys.argv[4] == "name"),
                      segment=int(sys.argv[5]) if len(sys.argv) > 5 else None,
                      gap_ms=float(sys.argv[6]) if len(sys.argv) > 6 else 0.5)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--pmc-traffic":
        # python bench.py --pmc-traffic FETCH.csv WRITE.csv N TILE OUT.json
        pmc_traffic(sys.argv[2], sys.argv[3], int(sys.argv[4]), sys.argv[5], sys.argv[6])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--grains", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed", help="comma list of fp64, mixed")
    ap.add_argument("--krylov", type=int, default=6, help="FGMRES restart around the MG cycle (0: plain MG)")
    ap.add_argument("--table", type=int, default=1, help="grain-table stiffness on level 0 for the headline mixed run")
    ap.add_argument("--rves", type=int, default=512, help="batched RVE count (0: skip the batched workload)")
    ap.add_argument("--plain-sample", type=int, default=24, help="RVEs solved one at a time by the plain FP64 solver")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="arrival window of the serving stream (0: skip serving)")
    ap.add_argument("--serve-wall", type=float, default=60.0, help="wall-time limit of the serving run")
    ap.add_argument("--serve-light", type=float, default=20.0,
                    help="arrival rate of a second, unsaturated serving run (0: skip it)")
    ap.add_argument("--serve-light-seconds", type=float, default=10.0, help="arrival window of the unsaturated run")
    ap.add_argument("--inclusion-n", type=int, default=64, help="cells per side of the 2D inclusion (0: skip it)")
    ap.add_argument("--only", default="poly,inclusion,rves,serve,amr",
                    help="comma list of the workloads to run (poly: the 256^3 headline; the oracle leg needs it)")
    ap.add_argument("--amr-n", type=int, default=128, help="base cells of the AMR crack nest (0: skip it)")
    ap.add_argument("--graphs", type=int, default=1,
                    help="also time the graph-replayed variants (0: skip them, e.g. under rocprofv3 --kernel-trace, "
                         "where a HIP graph launch has crashed the profiled process)")
    ap.add_argument("--planes", type=int, default=1,
                    help="also solve with per-node stiffness planes (0: only the grain-table solves, e.g. to profile "
                         "the headline)")
    ap.add_argument("--tile", default="",
                    help="fixed level-0 marching shape KZ/XCD/ROWS/PREFETCH (as the autotune keys), no autotuning "
                         "(the counter runs, whose serialized launches would mislead the autotuner)")
    ap.add_argument("--kz", type=int, default=0,
                    help="marching tile height of the 256^3 polycrystal hierarchy (0: autotune it before the timed region)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path, so --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)

    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K

    n = args.n
    nodes = (n + 1) ** 3
    poly = n > 0 and "poly" in args.only.split(",")
    tmx = t64 = t64t = tmp = tgr = st64 = stmx = kt = None
    umx = tuned = res_norms = disp_diff = probe_GBps = probe_var = fa64 = fa32 = ga64 = ga32 = latency = None
    dom = dom_ms = dom_bytes = achieved = traffic = traffic_tile = timed_tile = levels = plain_cycles = None
    peak, nc = 8000.0, 21
    if poly:
        log("bench: building the polycrystal workload n=%d grains=%d" % (n, args.grains))
        cfg = K.polycrystal(n, grains=args.grains)
        op = ela.Elastic(cfg["n"], ela.MM, mode=ela.ALLOC_BOTH)
        op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
        op.SetBC(cfg["bc_type"], cfg["bc_value"])
        op.SetKrylov(args.krylov)
        op.prepareForSolve()
        tuned = None
        if args.table and args.tile:  # a fixed marching shape
            kz_, xcd_, ry_, pf_ = (int(x) for x in args.tile.split("/"))
            for opt, val in ((ela.OPT_KZ, kz_), (ela.OPT_XCD, xcd_), (ela.OPT_ROWS, ry_), (ela.OPT_PREFETCH, pf_)):
                op.set_option(opt, val)
        elif args.table and args.kz:  # a fixed marching tile height (XCD-contiguous tile order)
            op.set_option(ela.OPT_KZ, args.kz)
            op.set_option(ela.OPT_XCD, 1)
        elif args.table:  # marching tile shape for this hierarchy, picked before the timed region
            op.SetStiffnessMode(1)
            tuned = {"%d/%d/%d/%d" % k: v for k, v in op.autotune(nrep=10).items()}
            op.SetStiffnessMode(0)
        b_host = op.bc_init()
        b = op.field(0, ela.F64).upload(b_host)
        u = op.field(0, ela.F64)
        nodes = (n + 1) ** 3
        log("bench: hierarchy %d MG levels, %.2f GB on device" % (op.levels(), op.bytes() / 1e9))

        modes = args.modes.split(",")
        t64 = st64 = u64 = None
        if "fp64" in modes and args.planes:  # plain FP64 GPU solver, per-node stiffness planes (the reference's storage)
            log("bench: fp64 solves (stiffness planes)")
            t64, per64, st64, rc64, _ = run_mode(op, b, u, ela.SOLVE_FP64, args.steps, args.warmup, args.tol)
            u64 = u.download()
            log("bench: fp64 %.3f s/solve iters=%d rel=%.2e rc=%d" % (t64 / args.steps, st64["iters"], st64["res"] / st64["res0"], rc64))
        tmp = stmp = None
        if "mixed" in modes and args.table and args.planes:  # mixed with planes, for the precision-only ratio
            log("bench: mixed solves (stiffness planes)")
            tmp, _, stmp, _, _ = run_mode(op, b, u, ela.SOLVE_MIXED, args.steps, args.warmup, args.tol)
            log("bench: mixed planes %.3f s/solve iters=%d" % (tmp / args.steps, stmp["iters"]))
        op.SetStiffnessMode(1 if args.table else 0)
        t64t = None
        if "fp64" in modes and args.table:
            log("bench: fp64 solves (grain table)")
            t64t, _, st64t, _, _ = run_mode(op, b, u, ela.SOLVE_FP64, args.steps, args.warmup, args.tol)
            log("bench: fp64 table %.3f s/solve iters=%d" % (t64t / args.steps, st64t["iters"]))
        # mixed precision (headline)
        log("bench: mixed solves (%s)" % ("grain table" if args.table else "stiffness planes"))
        tmx, permx, stmx, rcmx, kt = run_mode(op, b, u, ela.SOLVE_MIXED, args.steps, args.warmup, args.tol, timing=True)
        umx = u.download()
        log("bench: mixed %.3f s/solve iters=%d rel=%.2e rc=%d" % (tmx / args.steps, stmx["iters"], stmx["res"] / stmx["res0"], rcmx))
        disp_diff = float(np.abs(umx - u64).max() / np.abs(u64).max()) if u64 is not None else None
        # the mixed solution's residual in both norms (max-norm is the convergence test; u0 = 0, so r0 = b)
        r_mx, rinf_mx = op.residual(umx, b_host)
        res_norms = {"inf": rinf_mx / float(np.abs(b_host).max()),
                     "l2": float(np.linalg.norm(r_mx) / np.linalg.norm(b_host))}
        del r_mx
        # the same mixed solves with the cycles' launches on levels >= 1 replayed from HIP graphs (opt-in)
        tgr = None
        if args.table and args.graphs:
            op.set_option(ela.OPT_GRAPH, 1)
            tgr, _, stgr, _, _ = run_mode(op, b, u, ela.SOLVE_MIXED, args.steps, args.warmup, args.tol)
            op.set_option(ela.OPT_GRAPH, 0)
            log("bench: mixed, graph-replayed levels >= 1: %.3f s/solve iters=%d" % (tgr / args.steps, stgr["iters"]))
        op.SetStiffnessMode(0)
        # the plain MG iteration of the reference (per-node planes, FP64, no Krylov accelerator) once: its
        # cycle count is the oracle's at this size (the same algorithm), used by the CPU baseline
        plain_cycles = None
        if not args.no_cpu:
            op.SetKrylov(0)
            u.zero()
            op.prepareForSolve()
            rcp, stp = op.solve_device(b, u, args.tol, 0.0, ela.SOLVE_FP64, 400)
            plain_cycles = stp["iters"]
            op.SetKrylov(args.krylov)
            log("bench: plain MG FP64 (planes) %d cycles, rel %.2e" % (plain_cycles, stp["res"] / stp["res0"]))
        # HBM copy probe (also the byte-counter calibration kernel of the PMC runs)
        probe_bytes = 1 << 30
        probe_var = {v: 2 * probe_bytes / (op.ctx.probe_copy(probe_bytes, 5, v) * 1e-3) / 1e9 for v in range(7)}
        probe_GBps = max(probe_var.values())
        # what a device-wide barrier and a bare launch cost (the floor under the small levels' kernels)
        latency = {}
        for nb in (8, 64, 256):
            bar, launch = op.ctx.probe_barrier(nb, 2000)
            latency[str(nb)] = {"barrier_us": bar, "empty_launch_us": launch}

        # Fapply alone (FP64 and FP32), HIP events on the library stream
        fa64 = op.time_fapply(ela.F64, 20)
        fa32 = op.time_fapply(ela.F32, 20)
        ga64 = ga32 = None
        if args.table:  # the same operator with the grain-form stiffness (2-byte grain id per node)
            op.SetStiffnessMode(1)
            ga64 = op.time_fapply(ela.F64, 20)
            ga32 = op.time_fapply(ela.F32, 20)
            op.SetStiffnessMode(0)
        nc = 21
        bytes_fa64 = 8 * (3 + 3 + nc) * nodes
        bytes_fa32 = 4 * (3 + 3 + nc) * nodes
        # dominant kernel of the mixed solve: level-0 FP32 Jacobi sweep vs level-0 FP64 residual.
        # Jacobi bytes per node: x in, b, diagonal, x out (3 comps each) + the stiffness: 21 planes; or in
        # grain form x in, b, x out + a 2-byte grain id (the grain records and an interior node's diagonal,
        # which depends on its own record only, sit in LDS, read once per block), the diagonal field read
        # on the shell's nodes only
        shell = nodes - (n - 1) ** 3
        bytes_jac32 = (4 * (4 * 3 + nc)) * nodes if not args.table else (4 * 9 + 2) * nodes + 12 * shell
        # x, b, C in FP64 + scaled FP32 rhs out (plain MG) / A z: z, C in, w out (FGMRES)
        cbytes = 8 * nc if not args.table else 2
        bytes_res64 = ((8 * 6 + cbytes + 4 * 3) if not args.krylov else 8 * 6 + cbytes) * nodes
        tot_jac = kt["smooth0_ms"] * kt["smooth0_n"]
        tot_res = kt["resid64_ms"] * kt["resid64_n"]
        if tot_jac >= tot_res:
            dom = ("k_grain_sweep<float,0> (level-0 Jacobi sweep, grain form: shell + marching tiles)" if args.table
                   else "k_jacobi<float,MM,3,PlaneAcc> (level-0 Jacobi sweep)")
            dom_ms, dom_bytes = kt["smooth0_ms"], bytes_jac32
        else:
            dom = ("k_grain_sweep<double,%d> (level 0, grain form)" % (3 if args.krylov else 2)
                   if args.table else "k_residual<double,float,MM,3,PlaneAcc> (level 0)")
            dom_ms, dom_bytes = kt["resid64_ms"], bytes_res64
        peak = 8000.0
        achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
        traffic = traffic_tile = None
        tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
        if os.path.exists(tpath):  # HBM bytes per launch from the rocprofv3 --pmc runs, copy-calibrated
            with open(tpath) as fh:
                tr = json.load(fh)
            ent = tr.get("kernels", {}).get(dom)
            if ent and tr.get("n") == n:
                traffic = ent["fetch_bytes"] + ent["write_bytes"]
                traffic_tile = tr.get("tile")
        timed_tile = args.tile or (min(tuned, key=tuned.get) if tuned else None)

        levels = op.levels()
        op.close()
    keep = {"poly_full": umx} if poly else {}
    want = args.only.split(",")
    incl = inclusion_2d(args.inclusion_n, args.steps, keep) if args.inclusion_n > 0 and "inclusion" in want else None
    if not args.no_cpu and poly:
        small_solutions(args.cpu_n, 4, args.tol, keep)
    batched = (batched_rves(args.rves, args.plain_sample, args.tol, graphs=args.graphs) if args.rves > 0 and "rves" in want
               else None)
    serve = (serving(args.serve_seconds, args.serve_wall, args.tol) if args.serve_seconds > 0 and "serve" in want
             else None)
    if serve is not None and args.serve_light > 0:
        # the same stream at a rate one GPU keeps up with: latencies are service times, not queueing
        serve["unsaturated"] = serving(args.serve_light_seconds, args.serve_wall, args.tol, rate=args.serve_light,
                                       check=False)
    amr = amr_crack(args.amr_n, args.steps, args.tol, args.krylov, keep) if args.amr_n > 0 and "amr" in want else None

    cpu = checks = None
    if not args.no_cpu and poly:
        log("bench: oracle leg: cpu baseline (full solve at %d^3, set-up + 2 F-cycles at %d^3) and checks"
            % (args.cpu_n, n))
        leg = cpu_and_checks(n, args.cpu_n, plain_cycles, keep)
        ci, sm, fu = leg["cpu"], leg["small"], leg["full"]
        cpu = {"value": leg["value"], "unit": "solves/s", "cores": leg["threads"], "kind": "port",
               "cpu_model": ci["model"], "nproc": ci["nproc"], "affinity": ci["affinity"],
               "single_thread_cycle_s_at_%d" % args.cpu_n: sm["single_thread_cycle_s"],
               "sample": "the FP64 oracle (C++ restatement of the reference algorithm: plain MG F-cycles, Jacobi "
                         "2/3 x2, dense bottom; OpenMP, %d threads) on the same workload: at %d^3 its set-up "
                         "(%.1f s) and 2 F-cycles (%.2f s each) were timed and the solve is set-up + %d cycles (the "
                         "GPU's plain-MG FP64 cycle count of the same iteration at %d^3) = %.0f s; at %d^3 a full "
                         "solve took %.1f s (%d cycles), one F-cycle on one thread %.2f s"
                         % (leg["threads"], n, fu["setup_s"], fu["cycle_s"], plain_cycles, n, fu["solve_s"],
                            args.cpu_n, sm["seconds"], sm["cycles"], sm["single_thread_cycle_s"])}
        checks = leg["per_config_error"]

    if not poly:  # a partial run (profiling one workload): no headline value
        print(json.dumps({"metric": "solves/s at 1e-10 rel residual (mixed FP32/FP64), vs plain FP64 GPU",
                          "value": None, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
                          "only": args.only, "inclusion_2d": incl, "batched_rves": batched, "serving": serve,
                          "amr_crack": amr}), flush=True)
        return
    out = {
        "metric": "solves/s at 1e-10 rel residual (mixed FP32/FP64), vs plain FP64 GPU",
        "value": args.steps / tmx,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "3D single level %d^3 cells, cubic Voronoi %d grains (Haar orientations, seed %d), "
                               "uniaxial z strain 0.01, tol_rel %.0e" % (n, args.grains, K.SEED, args.tol),
                   "nodes": nodes, "mg_levels": levels,
                   "solver": "FGMRES(%d) + MG F-cycle" % args.krylov if args.krylov else "MG F-cycles"},
        "fp64_solves_per_s": args.steps / t64t if t64t else (args.steps / t64 if t64 else None),
        "mixed_over_fp64": (t64t / tmx) if t64t else (t64 / tmx if t64 else None),
        "mixed_over_fp64_note": "same stiffness storage (grain table) for both precisions; the per-node-plane "
                                "FP64 solver is in variants_solves_per_s",
        "mixed_table_over_fp64_planes": t64 / tmx if t64 else None,
        "stiffness": "grain table on level 0 (16-bit grain id + per-grain records)" if args.table else "per-node planes",
        "autotune_ms_per_sweep": tuned,
        "variants_solves_per_s": {"fp64_planes": args.steps / t64 if t64 else None,
                                  "mixed_planes": args.steps / tmp if tmp else None,
                                  "fp64_table": args.steps / t64t if t64t else None,
                                  "mixed_table": args.steps / tmx if args.table else None,
                                  "mixed_table_graph": args.steps / tgr if tgr else None},
        "iters": {"fp64": st64["iters"] if st64 else None, "mixed": stmx["iters"]},
        "rel_residual": {"fp64": st64["res"] / st64["res0"] if st64 else None, "mixed": stmx["res"] / stmx["res0"]},
        "hbm_copy_probe_GBps": probe_GBps,
        "hbm_copy_probe_variants_GBps": probe_var,
        "launch_latency_by_blocks": latency,
        "rel_disp_diff_mixed_vs_fp64": disp_diff,
        "mixed_rel_residual_norms": res_norms,
        "fapply": {"fp64_ms": fa64, "fp64_GBps": bytes_fa64 / (fa64 * 1e-3) / 1e9,
                   "fp32_ms": fa32, "fp32_GBps": bytes_fa32 / (fa32 * 1e-3) / 1e9,
                   "bytes_per_node_fp64": 8 * (3 + 3 + nc),
                   "grain_form": None if ga64 is None else {
                       "fp64_ms": ga64, "fp64_GBps": (8 * 6 + 2) * nodes / (ga64 * 1e-3) / 1e9,
                       "fp32_ms": ga32, "fp32_GBps": (4 * 6 + 2) * nodes / (ga32 * 1e-3) / 1e9,
                       "bytes_per_node_fp64": 8 * 6 + 2, "bytes_per_node_fp32": 4 * 6 + 2,
                       "fp64_speedup_over_planes": fa64 / ga64}},
        "kernel_times": kt,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": peak, "unit": "GB/s", "frac": achieved / peak,
                     "tile": timed_tile, "traffic_tile": traffic_tile,
                     "traffic": traffic, "kernel": dom, "kernel_ms": dom_ms, "bytes_per_launch": dom_bytes},
        "cpu_baseline": cpu,
        "per_config_error": checks,
        "inclusion_2d": incl,
        "batched_rves": batched,
        "serving": serve,
        "amr_crack": amr,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
