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


def cpu_baseline(n):
    """The oracle (FP64 CPU restatement of the reference algorithm: plain MG F-cycles, Jacobi smoothing)
    solving the same workload at n^3 to 1e-10; the solve time is scaled to 256^3 by node count (MG cycle
    counts do not depend on the size)."""
    from anisotropicfracture_amd import configs as K
    from oracle import oracle as O
    O.build()
    cfg = K.polycrystal(n)
    gid = K.voronoi_grain_ids(cfg["seeds"], cfg["n"])
    tab = np.stack([O.cubic_define(cfg["C11"], cfg["C12"], cfg["C44"], R) for R in cfg["R"]])
    C = np.moveaxis(tab[gid], -1, 0).copy()
    g = O.Grid(cfg["n"])
    b = O.bc_init(g, cfg["bc_type"], cfg["bc_value"], g.zeros())
    t0 = time.perf_counter()
    _, st, _ = O.mg_solve(g, O.MM, C, cfg["bc_type"], b, tol_rel=1e-10, max_iter=400)
    t = time.perf_counter() - t0
    return t, st["iters"], st["converged"]


def batched_rves(count, sample, tol):
    """Batched RVEs: `count` independent RVEs (sides 32/48/64/96, 8-32 grains, fixed-corner uniaxial z
    strain 0.01).  Batched mixed and batched FP64 solve every RVE, one launch per cycle kernel over all
    of them; the plain FP64 GPU solver (ela_solve on one hierarchy per size, one RVE at a time, the same
    MG iteration) solves the first `sample` RVEs.  Timed per RVE: stiffness evaluation on the device,
    set-up (coarse stiffness, diagonals, coarsest inverse) and the solve."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    cfg = K.rve_batch(count)
    out = {"rves": count, "nodes": int(sum((m["n"] + 1) ** 3 for m in cfg))}
    cycles = {}

    def load(b, rs, ms):
        for r, m in zip(rs, ms):
            r.SetModelVoronoiCubic(m["seeds"], K.VORONOI_Q, m["R"].reshape(-1, 9), K.C11, K.C12, K.C44)
            r.SetRhsAffine(m["G"])

    for mode, key in ((ela.SOLVE_MIXED, "mixed"), (ela.SOLVE_FP64, "fp64")):
        b = ela.Batch(mode)
        # untimed warm-up batch: first launches of every kernel, and the marching tile shape picked on
        # a launch large enough to fill the GPU
        nw = min(64, count)
        warm = [b.rve(m["n"], max_grains=32) for m in cfg[:nw]]
        load(b, warm, cfg[:nw])
        b.prepare(warm)
        tuned = {"%d/%d" % k: v for k, v in b.autotune(warm, nrep=3).items()}
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
        if key == "mixed":
            sm = b.smoother_times()
            ent["smoother"] = {"ms": sm["ms"], "launches": sm["launches"], "node_sweeps": sm["nodes"],
                               "Gnodes_per_s": sm["nodes"] / (sm["ms"] * 1e-3) / 1e9}
        out["batched_" + key] = ent
        log("bench: batched %s %.1f solves/s (%.2f s, cycles %.1f)" % (key, count / t, t, it.mean()))
        b.close()
    # plain FP64 GPU solver, one RVE at a time (hierarchies reused per size)
    ops, tot, its = {}, 0.0, []
    for m in cfg[:sample]:
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
        tot += time.perf_counter() - t0
        its.append(st["iters"])
    for op, _, _ in ops.values():
        op.close()
    out["plain_fp64"] = {"solves_per_s": sample / tot, "sample": sample, "cycles_mean": float(np.mean(its)),
                         "cycles_same_as_batched_fp64": bool(np.array_equal(its, cycles["fp64"][:sample]))}
    out["batched_mixed_over_plain_fp64"] = out["batched_mixed"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    out["batched_fp64_over_plain_fp64"] = out["batched_fp64"]["solves_per_s"] / out["plain_fp64"]["solves_per_s"]
    log("bench: plain fp64 %.1f solves/s (sample %d); batched mixed / plain fp64 = %.2f"
        % (sample / tot, sample, out["batched_mixed_over_plain_fp64"]))
    return out


def amr_crack(n, steps, tol, krylov):
    """The mode-I crack on a 3-level AMR nest (128^3 base) (composite solve), mixed and FP64.

    A step is prepareForSolve + one composite solve from u = 0 with b and the model resident."""
    import anisotropicfracture_amd as ela
    from anisotropicfracture_amd import configs as K
    cfg = K.crack_nest(n)
    out = {"base": n, "boxes": [[list(lo), list(hi)] for lo, hi in cfg["boxes"]], "g": list(cfg["g"]),
           "width": cfg["width"]}
    sols = {}
    for key, mode in (("fp64", ela.SOLVE_FP64), ("mixed", ela.SOLVE_MIXED)):
        a = ela.ElasticAMR(cfg["n"], cfg["boxes"], ela.ISO, mode=mode, coarse_min=4)
        a.SetBC(cfg["bc_type"], cfg["bc_value"])
        a.SetKrylov(krylov)
        nodes = []
        for l in range(a.nlev):
            X = a.coords(l)
            a.SetModel(l, K.crack_nest_model(cfg, X[0], X[1], X[2]))
            a.upload(a.B, l, a.bc_init(l))
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
           "requested_Mnodes": float(sum((q["n_requested"] + 1) ** 3 for _, q in done) / 1e6) if done else None,
           "not_served_by_limit": int(sum(1 for r in res if r["status"] == -1)),
           "time_breakdown_s": summary}
    log("bench: serving at %.0f/s: %d/%d requests in %.1f s: %.1f solves/s p50 %.0f ms p99 %.0f ms"
        % (rate, len(done), len(reqs), wall, out["solves_per_s"], out["p50_ms"] or -1, out["p99_ms"] or -1))
    return out


def main():
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
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the serving stream (0: skip serving)")
    ap.add_argument("--serve-wall", type=float, default=40.0, help="wall-time limit of the serving run")
    ap.add_argument("--serve-light", type=float, default=20.0,
                    help="arrival rate of a second, unsaturated serving run (0: skip it)")
    ap.add_argument("--amr-n", type=int, default=128, help="base cells of the AMR crack nest (0: skip it)")
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
    log("bench: building the polycrystal workload n=%d grains=%d" % (n, args.grains))
    cfg = K.polycrystal(n, grains=args.grains)
    op = ela.Elastic(cfg["n"], ela.MM, mode=ela.ALLOC_BOTH)
    op.SetModelVoronoiCubic(cfg["seeds"], K.VORONOI_Q, cfg["R"], cfg["C11"], cfg["C12"], cfg["C44"])
    op.SetBC(cfg["bc_type"], cfg["bc_value"])
    op.SetKrylov(args.krylov)
    op.prepareForSolve()
    tuned = None
    if args.table and args.kz:  # a fixed marching tile height (XCD-contiguous tile order)
        op.set_option(ela.OPT_KZ, args.kz)
        op.set_option(ela.OPT_XCD, 1)
    elif args.table:  # marching tile shape for this hierarchy, picked before the timed region
        op.SetStiffnessMode(1)
        tuned = {"%d/%d" % k: v for k, v in op.autotune(nrep=10).items()}
        op.SetStiffnessMode(0)
    b_host = op.bc_init()
    b = op.field(0, ela.F64).upload(b_host)
    u = op.field(0, ela.F64)
    nodes = (n + 1) ** 3
    log("bench: hierarchy %d MG levels, %.2f GB on device" % (op.levels(), op.bytes() / 1e9))

    modes = args.modes.split(",")
    t64 = st64 = u64 = None
    if "fp64" in modes:  # plain FP64 GPU solver, per-node stiffness planes (the reference's storage)
        log("bench: fp64 solves (stiffness planes)")
        t64, per64, st64, rc64, _ = run_mode(op, b, u, ela.SOLVE_FP64, args.steps, args.warmup, args.tol)
        u64 = u.download()
        log("bench: fp64 %.3f s/solve iters=%d rel=%.2e rc=%d" % (t64 / args.steps, st64["iters"], st64["res"] / st64["res0"], rc64))
    tmp = stmp = None
    if "mixed" in modes and args.table:  # mixed with planes, for the precision-only ratio
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
    if args.table:
        op.set_option(ela.OPT_GRAPH, 1)
        tgr, _, stgr, _, _ = run_mode(op, b, u, ela.SOLVE_MIXED, args.steps, args.warmup, args.tol)
        op.set_option(ela.OPT_GRAPH, 0)
        log("bench: mixed, graph-replayed levels >= 1: %.3f s/solve iters=%d" % (tgr / args.steps, stgr["iters"]))
    op.SetStiffnessMode(0)
    # HBM copy probe (also the byte-counter calibration kernel of the PMC runs)
    probe_bytes = 1 << 30
    probe_ms = op.ctx.probe_copy(probe_bytes, 5)
    probe_GBps = 2 * probe_bytes / (probe_ms * 1e-3) / 1e9

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
    # Jacobi bytes per node: x in, b, diagonal, x out (3 comps each) + the stiffness: 21 planes, or a
    # 2-byte grain id (the grain records sit in LDS / L1, read once per block)
    bytes_jac32 = (4 * (4 * 3 + nc) if not args.table else 4 * 12 + 2) * nodes
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
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):  # HBM bytes per launch from the rocprofv3 --pmc runs, copy-calibrated
        with open(tpath) as fh:
            tr = json.load(fh)
        ent = tr.get("kernels", {}).get(dom)
        if ent and tr.get("n") == n:
            traffic = ent["fetch_bytes"] + ent["write_bytes"]

    levels = op.levels()
    op.close()
    batched = batched_rves(args.rves, args.plain_sample, args.tol) if args.rves > 0 else None
    serve = serving(args.serve_seconds, args.serve_wall, args.tol) if args.serve_seconds > 0 else None
    if serve is not None and args.serve_light > 0:
        # the same stream at a rate one GPU keeps up with: latencies are service times, not queueing
        serve["unsaturated"] = serving(args.serve_seconds, args.serve_wall, args.tol, rate=args.serve_light,
                                       check=False)
    amr = amr_crack(args.amr_n, args.steps, args.tol, args.krylov) if args.amr_n > 0 else None

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log("bench: cpu baseline (oracle, %d threads) full solve at %d^3" % (threads, args.cpu_n))
        tcpu, cyc, conv = cpu_baseline(args.cpu_n)
        scale = nodes / float((args.cpu_n + 1) ** 3)
        t_solve = tcpu * scale
        cpu = {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": "one full solve to 1e-10 of the same workload at %d^3 by the FP64 oracle (plain MG "
                         "F-cycles, %d cycles, %.1f s, converged=%s), scaled by node count to %d^3"
                         % (args.cpu_n, cyc, tcpu, conv, n)}

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
        "fp64_solves_per_s": args.steps / t64 if t64 else None,
        "mixed_over_fp64": t64 / tmx if t64 else None,
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
                     "traffic": traffic, "kernel": dom, "kernel_ms": dom_ms, "bytes_per_launch": dom_bytes},
        "cpu_baseline": cpu,
        "batched_rves": batched,
        "serving": serve,
        "amr_crack": amr,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
